// Batched eigen-decomposition of general complex FP64 matrices (the P*Q
// eigenproblem of a patterned layer, khepri/alternative.py:172).
//
// One workgroup per matrix, four stages:
//   1. Householder reduction to upper Hessenberg form H = Z^H A Z (Z explicit);
//   2. implicit single-shift complex QR (Wilkinson shift, exceptional shifts
//      every 10 sweeps, Ahues-Tisseur deflation test) to the Schur form
//      T = Z^H A Z, accumulating Z;
//   3. eigenvectors of the triangular T by back substitution (column-parallel
//      over eigenvectors, row-sequential), tiny pivots replaced by
//      max(ulp*|lambda|_1, smallnum) as LAPACK does;
//   4. (separate launches) W = Z X, columns normalised to unit 2-norm with the
//      largest component real -- the zgeev convention.
// Each QR rotation touches disjoint row/column/Z element sets, so a rotation
// is one barrier; the inputs of the next rotation travel through LDS.
#include "khx_internal.h"
#include "khx_prof.h"

namespace khx {

constexpr int EIG_THREADS = 256;
constexpr int EIG_WAVES = EIG_THREADS / 64;
constexpr double ULP = 2.220446049250313e-16;
constexpr double SAFMIN = 2.2250738585072014e-308;

__device__ __forceinline__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  v = wave_sum(v);
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < EIG_WAVES; ++w) t += red[w];
  __syncthreads();
  return t;
}

__device__ __forceinline__ int block_max_int(int v, int* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  if (lane == 0) red[wid] = v;
  __syncthreads();
  int t = red[0];
#pragma unroll
  for (int w = 1; w < EIG_WAVES; ++w) t = max(t, red[w]);
  __syncthreads();
  return t;
}

// Complex Givens rotation G = [[c, s], [-conj(s), c]] with G [x; y] = [r; 0].
__device__ __forceinline__ void givens(cd x, cd y, double& c, cd& s, cd& r) {
  const double ax = absc(x), ay = absc(y);
  if (ay == 0.0) {
    c = 1.0;
    s = mkc(0.0, 0.0);
    r = x;
    return;
  }
  if (ax == 0.0) {
    c = 0.0;
    s = (1.0 / ay) * conjc(y);
    r = mkc(ay, 0.0);
    return;
  }
  const double nr = hypot(ax, ay);
  c = ax / nr;
  const cd ph = (1.0 / ax) * x;  // x/|x|
  s = (1.0 / nr) * (ph * conjc(y));
  r = nr * ph;
}

// Negligible sub-diagonal H[k][k-1] (conventional test + Ahues & Tisseur).
__device__ __forceinline__ bool negligible(const cd* H, int n, int k, double smlnum) {
  const cd hkk1 = H[(long long)k * n + k - 1];
  const double h = cabs1(hkk1);
  if (h <= smlnum) return true;
  const cd a = H[(long long)(k - 1) * n + k - 1];
  const cd d = H[(long long)k * n + k];
  double tst = cabs1(a) + cabs1(d);
  if (tst == 0.0) return false;
  if (h > ULP * tst) return false;
  const double up = cabs1(H[(long long)(k - 1) * n + k]);
  const double ab = fmax(h, up), ba = fmin(h, up);
  const double dd = cabs1(a - d), ad = cabs1(d);
  const double aa = fmax(ad, dd), bb = fmin(ad, dd);
  const double s = aa + ab;
  return ba * (ab / s) <= fmax(smlnum, ULP * (bb * (aa / s)));
}

// Stages 1-3.  A (n x n) becomes T; Z gets the Schur vectors; X the
// eigenvectors of T (upper triangular, unit diagonal); lam the eigenvalues.
__global__ __launch_bounds__(EIG_THREADS) void zeig_schur_kernel(int n, cd* A, long long sA, cd* Zg,
                                                                 cd* Xg, long long sW, cd* lam,
                                                                 long long sL, int* info) {
  extern __shared__ cd sm[];
  cd* u = sm;            // n   Householder vector
  cd* wv = u + n;        // n   u^H H  (left product)
  cd* vh = wv + n;       // n   H u    (right product)
  cd* vz = vh + n;       // n   Z u
  cd* part = vz + n;     // EIG_WAVES * n partial sums
  __shared__ double red[EIG_WAVES];
  __shared__ int redi[EIG_WAVES];
  __shared__ cd s_xy[2][2];  // double-buffered inputs of the next rotation

  const long long item = blockIdx.x;
  cd* H = A + item * sA;
  cd* Z = Zg + item * sW;
  cd* X = Xg + item * sW;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const double smlnum = SAFMIN * ((double)n / ULP);

  // Z = I
  for (int i = wid; i < n; i += EIG_WAVES)
    for (int j = lane; j < n; j += 64) Z[(long long)i * n + j] = mkc(i == j ? 1.0 : 0.0, 0.0);
  __syncthreads();

  // ------------------------------------------------------------ 1. Hessenberg
  for (int k = 0; k + 2 < n; ++k) {
    const int L = n - k - 1;  // reflector length (rows k+1..n-1)
    double part2 = 0.0;
    for (int i = k + 2 + tid; i < n; i += EIG_THREADS) part2 += abs2c(H[(long long)i * n + k]);
    const double rest2 = block_sum(part2, red);
    if (rest2 == 0.0) continue;  // uniform: every thread sees the same sum
    const cd alpha = H[(long long)(k + 1) * n + k];
    const double aa = absc(alpha);
    const double xn = sqrt(aa * aa + rest2);
    const cd ph = aa != 0.0 ? (1.0 / aa) * alpha : mkc(1.0, 0.0);
    const double sc = 1.0 / sqrt(xn * (xn + aa));
    for (int i = tid; i < L; i += EIG_THREADS) {
      cd x = H[(long long)(k + 1 + i) * n + k];
      if (i == 0) x = x + xn * ph;
      u[i] = sc * x;
    }
    __syncthreads();
    // left: w_j = sum_i conj(u_i) H[k+1+i][j], j > k
    for (int j = k + 1 + lane; j < n; j += 64) {
      cd acc = mkc(0.0, 0.0);
      for (int i = wid; i < L; i += EIG_WAVES) acc = fmac_cj(acc, u[i], H[(long long)(k + 1 + i) * n + j]);
      part[wid * n + j] = acc;
    }
    __syncthreads();
    for (int j = k + 1 + tid; j < n; j += EIG_THREADS) {
      cd acc = part[j];
#pragma unroll
      for (int w = 1; w < EIG_WAVES; ++w) acc = acc + part[w * n + j];
      wv[j] = acc;
    }
    __syncthreads();
    for (int i = wid; i < L; i += EIG_WAVES) {
      const cd ui = u[i];
      cd* row = H + (long long)(k + 1 + i) * n;
      for (int j = k + 1 + lane; j < n; j += 64) row[j] = fmsc(row[j], ui, wv[j]);
      if (lane == 0) row[k] = (i == 0) ? -(xn * ph) : mkc(0.0, 0.0);
    }
    __syncthreads();
    // right: v_i = sum_j H[i][k+1+j] u_j, same for Z
    for (int i = wid; i < n; i += EIG_WAVES) {
      const cd* hr = H + (long long)i * n + k + 1;
      const cd* zr = Z + (long long)i * n + k + 1;
      cd ah = mkc(0.0, 0.0), az = mkc(0.0, 0.0);
      for (int j = lane; j < L; j += 64) {
        ah = fmac(ah, hr[j], u[j]);
        az = fmac(az, zr[j], u[j]);
      }
      ah.x = wave_sum(ah.x);
      ah.y = wave_sum(ah.y);
      az.x = wave_sum(az.x);
      az.y = wave_sum(az.y);
      if (lane == 0) {
        vh[i] = ah;
        vz[i] = az;
      }
    }
    __syncthreads();
    for (int i = wid; i < n; i += EIG_WAVES) {
      cd* hr = H + (long long)i * n + k + 1;
      cd* zr = Z + (long long)i * n + k + 1;
      const cd a = vh[i], b = vz[i];
      for (int j = lane; j < L; j += 64) {
        const cd cu = conjc(u[j]);
        hr[j] = fmsc(hr[j], a, cu);
        zr[j] = fmsc(zr[j], b, cu);
      }
    }
    __syncthreads();
  }

  // ------------------------------------------------------------ 2. Schur form
  int ihi = n - 1, its = 0, itot = 0, fail = 0;
  const int itmax = 30 * max(10, n);
  while (ihi >= 1) {
    int lk = 0;
    for (int k = 1 + tid; k <= ihi; k += EIG_THREADS) {
      if (negligible(H, n, k, smlnum)) {
        H[(long long)k * n + k - 1] = mkc(0.0, 0.0);
        lk = max(lk, k);
      }
    }
    const int l = block_max_int(lk, redi);
    if (l == ihi) {
      --ihi;
      its = 0;
      continue;
    }
    if (++itot > itmax) {
      fail = ihi + 1;
      break;
    }
    ++its;
    // shift
    cd shift;
    if (its % 20 == 10) {  // exceptional shifts break cycles
      shift = H[(long long)l * n + l] + mkc(0.75 * fabs(H[(long long)(l + 1) * n + l].x), 0.0);
    } else if (its % 20 == 0) {
      shift = H[(long long)ihi * n + ihi] + mkc(0.75 * fabs(H[(long long)ihi * n + ihi - 1].x), 0.0);
    } else {
      const cd a = H[(long long)(ihi - 1) * n + ihi - 1], b = H[(long long)(ihi - 1) * n + ihi];
      const cd c = H[(long long)ihi * n + ihi - 1], d = H[(long long)ihi * n + ihi];
      const cd bc = b * c;
      if (bc.x == 0.0 && bc.y == 0.0) {
        shift = d;
      } else {
        const cd tr = 0.5 * (a - d);
        const cd disc = sqrtc(fmac(bc, tr, tr));
        const cd p1 = tr + disc, p2 = tr - disc;
        const cd den = abs2c(p1) >= abs2c(p2) ? p1 : p2;
        shift = (den.x == 0.0 && den.y == 0.0) ? d : d - divc(bc, den);
      }
    }
    cd x = H[(long long)l * n + l] - shift;
    cd y = H[(long long)(l + 1) * n + l];
    for (int k = l; k < ihi; ++k) {
      cd* nxt = s_xy[(k + 1) & 1];
      if (k > l) {
        x = s_xy[k & 1][0];
        y = s_xy[k & 1][1];
      }
      double c;
      cd s, r;
      givens(x, y, c, s, r);
      const cd sc = conjc(s);
      const int nrow = n - k - 2;                        // row part: columns k+2..n-1
      const int ncol = k + ((k + 2 <= ihi) ? 1 : 0);     // col part: rows 0..k-1 (+ row k+2)
      const int total = 1 + nrow + ncol + n;
      for (int e = tid; e < total; e += EIG_THREADS) {
        if (e == 0) {
          // 2x2 diagonal block G B G^H, and the bulge column k-1
          if (k > l) {
            H[(long long)k * n + k - 1] = r;
            H[(long long)(k + 1) * n + k - 1] = mkc(0.0, 0.0);
          }
          cd* p00 = H + (long long)k * n + k;
          cd* p10 = H + (long long)(k + 1) * n + k;
          const cd b00 = p00[0], b01 = p00[1], b10 = p10[0], b11 = p10[1];
          // rows: G B
          const cd r00 = fmac(c * b00, s, b10), r01 = fmac(c * b01, s, b11);
          const cd r10 = fmsc(c * b10, sc, b00), r11 = fmsc(c * b11, sc, b01);
          // columns: (G B) G^H
          const cd n00 = fmac(c * r00, sc, r01), n01 = fmsc(c * r01, s, r00);
          const cd n10 = fmac(c * r10, sc, r11), n11 = fmsc(c * r11, s, r10);
          p00[0] = n00;
          p00[1] = n01;
          p10[0] = n10;
          p10[1] = n11;
          nxt[0] = n10;
        } else if (e <= nrow) {
          const int j = k + 1 + e;  // k+2 ..
          cd* pk = H + (long long)k * n + j;
          cd* pk1 = H + (long long)(k + 1) * n + j;
          const cd hk = *pk, hk1 = *pk1;
          *pk = fmac(c * hk, s, hk1);
          *pk1 = fmsc(c * hk1, sc, hk);
        } else if (e <= nrow + ncol) {
          const int q = e - nrow - 1;
          const int i = (q < k) ? q : k + 2;
          cd* pr = H + (long long)i * n + k;
          const cd hk = pr[0], hk1 = pr[1];
          const cd nk = fmac(c * hk, sc, hk1);
          pr[0] = nk;
          pr[1] = fmsc(c * hk1, s, hk);
          if (i == k + 2) nxt[1] = nk;
        } else {
          const int i = e - nrow - ncol - 1;
          cd* pr = Z + (long long)i * n + k;
          const cd zk = pr[0], zk1 = pr[1];
          pr[0] = fmac(c * zk, sc, zk1);
          pr[1] = fmsc(c * zk1, s, zk);
        }
      }
      __syncthreads();
    }
  }
  if (tid == 0 && info) info[item] = fail;

  // eigenvalues
  cd* lm = lam + item * sL;
  for (int i = tid; i < n; i += EIG_THREADS) lm[i] = H[(long long)i * n + i];

  // ------------------------------------------------------------ 3. eigenvectors of T
  // X[i][k] = -T[i][k] (i < k), 1 (i == k), 0 (i > k)
  for (int i = wid; i < n; i += EIG_WAVES)
    for (int k = lane; k < n; k += 64) {
      const cd t = H[(long long)i * n + k];
      X[(long long)i * n + k] = (k > i) ? -t : mkc(k == i ? 1.0 : 0.0, 0.0);
    }
  __syncthreads();
  cd* xr = u;  // reuse: current row of solved components
  for (int j = n - 1; j >= 0; --j) {
    const cd tjj = H[(long long)j * n + j];
    for (int k = j + 1 + tid; k < n; k += EIG_THREADS) {
      const cd lk = H[(long long)k * n + k];
      cd d = tjj - lk;
      const double smin = fmax(ULP * cabs1(lk), smlnum);
      if (cabs1(d) < smin) d = mkc(smin, 0.0);
      const cd v = divc(X[(long long)j * n + k], d);
      X[(long long)j * n + k] = v;
      xr[k] = v;
    }
    __syncthreads();
    for (int i = wid; i < j; i += EIG_WAVES) {
      const cd tij = H[(long long)i * n + j];
      cd* row = X + (long long)i * n;
      for (int k = j + 1 + lane; k < n; k += 64) row[k] = fmsc(row[k], tij, xr[k]);
    }
    __syncthreads();
  }
}

// Column normalisation of W: unit 2-norm, largest |component| made real.
__global__ __launch_bounds__(256) void znormalize_cols_kernel(int n, cd* W, long long sW) {
  const long long item = blockIdx.x;
  cd* w = W + item * sW;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    double ss = 0.0, best = -1.0;
    int bi = 0;
    for (int i = 0; i < n; ++i) {
      const double a2 = abs2c(w[(long long)i * n + k]);
      ss += a2;
      if (a2 > best) {
        best = a2;
        bi = i;
      }
    }
    const double inv = 1.0 / sqrt(ss);
    const cd vb = inv * w[(long long)bi * n + k];
    const double ab = absc(vb);
    const cd rot = ab > 0.0 ? (1.0 / ab) * conjc(vb) : mkc(1.0, 0.0);
    const cd f = inv * rot;
    for (int i = 0; i < n; ++i) w[(long long)i * n + k] = f * w[(long long)i * n + k];
    w[(long long)bi * n + k].y = 0.0;
  }
}

size_t zeig_workspace_bytes(int n, int batch) {
  return 2ull * (size_t)n * n * sizeof(cd) * (size_t)batch;  // Z and X
}

int zeig(hipStream_t s, int n, Mat A, cd* lam, long long lstride, Mat W, int batch, int* info, void* ws) {
  if (batch <= 0) return 0;
  KHX_REQUIRE(n >= 1, "zeig: n must be positive");
  KHX_REQUIRE(A.ld == n && W.ld == n && A.stride != 0 && W.stride != 0,
              "zeig: matrices must be dense (ld == n) and not broadcast");
  cd* Z = static_cast<cd*>(ws);
  cd* X = Z + (size_t)n * n * batch;
  const long long sZ = (long long)n * n;
  const size_t lds = (size_t)(4 + EIG_WAVES) * n * sizeof(cd);
  KHX_REQUIRE(lds <= 64 * 1024, "zeig: n too large for the LDS vectors");
  // Z and X use the W stride layout (n*n per item) in the workspace
  {
  ProfScope ps("zeig_schur_kernel", s);
  zeig_schur_kernel<<<dim3(batch), dim3(EIG_THREADS), lds, s>>>(n, A.p, A.stride, Z, X, sZ, lam,
                                                               lstride, info);
  KHX_LAUNCH_CHECK();
  }
  // W = Z X
  KHX_TRY(zgemm(s, n, n, n, mkc(1.0, 0.0), Mat{Z, n, sZ}, Mat{X, n, sZ}, mkc(0.0, 0.0), W, batch));
  znormalize_cols_kernel<<<dim3(batch), dim3(256), 0, s>>>(n, W.p, W.stride);
  KHX_LAUNCH_CHECK();
  return 0;
}

}  // namespace khx
