// Band-structure tool (khepri/eigentricks.py:5-55): Bloch modes of a period
// from its S-matrix.
//
// S is a (4ng x 4ng) row-major matrix per item in the (+, -) block layout:
// p = rows/columns 0:n, m = n:2n with n = 2ng.  The Bloch factors w and modes
// v = (c+, c-) at the left face solve the pencil Sl v = w Sr v with
//   Sl = [[Spp, 0], [Smp, -I]],  Sr = [[I, -Spm], [0, -Smm]].
//
// The reference hands the pencil to a QZ solver.  Here it goes through the
// Cayley transform C = (Sl - s Sr)^-1 (Sl + s Sr) with a shift s on the unit
// circle: C has the same eigenvectors and the eigenvalues c = (w + s)/(w - s),
// so w = s (c + 1)/(c - 1).  Sl - s Sr has bounded entries (those of S) and is
// singular only where s is itself a Bloch factor, so unlike Sr^-1 Sl (which
// holds Smm^-1, exponentially large for thick evanescent periods) the
// propagating, on-shell factors |w| = 1 keep full accuracy; they map to the
// imaginary axis, c = i cot((arg w - arg s)/2).  C goes through the batched
// LU solve and the batched complex eigensolver of the layer solves.
//
// The determinant det(Spp - Spm Smm^-1 Smp) det(Smm) (khepri/eigentricks.py:24-28)
// is taken from two LU factorisations as sign * exp(sum log|u_ii|), the way
// numpy's det forms it.
#include "../../include/khx.h"
#include "khx_internal.h"

namespace khx {

// Sl, Sr (batch, 2n, 2n) from S (khepri/eigentricks.py:5-22).
__global__ __launch_bounds__(256) void splitlr_kernel(int n, const cd* __restrict__ S, cd* __restrict__ Sl,
                                                      cd* __restrict__ Sr, long long total) {
  const int n2 = 2 * n;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(e % n2), r = (int)((e / n2) % n2);
    const bool rp = r < n, cp = c < n;
    const cd s = S[e];
    const cd zero = mkc(0.0, 0.0);
    const cd eye = mkc(r == c ? 1.0 : 0.0, 0.0);
    Sl[e] = cp ? s : (rp ? zero : -eye);     // [[Spp, 0], [Smp, -I]]
    Sr[e] = cp ? (rp ? eye : zero) : -s;     // [[I, -Spm], [0, -Smm]]
  }
}

__global__ __launch_bounds__(256) void fill_int_kernel(int* __restrict__ p, int count, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

// flag[b] = 1 when S[b] holds a NaN (the reference then returns None).
__global__ __launch_bounds__(256) void nan_flag_kernel(const cd* __restrict__ S, long long per_item, long long total,
                                                       int* __restrict__ flag) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const cd s = S[e];
    if (isnan(s.x) || isnan(s.y)) atomicOr(flag + e / per_item, 1);
  }
}

// A = Sl - s Sr, B = Sl + s Sr; an item flagged NaN gets A = I, B = 0 (C = 0).
__global__ __launch_bounds__(256) void cayley_kernel(int n, const cd* __restrict__ S, cd sg,
                                                     const int* __restrict__ flag, cd* __restrict__ A,
                                                     cd* __restrict__ B, long long total) {
  const int n2 = 2 * n;
  const long long per = (long long)n2 * n2;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(e % n2), r = (int)((e / n2) % n2);
    const bool rp = r < n, cp = c < n;
    const double dg = r == c ? 1.0 : 0.0;
    const cd s = S[e];
    cd l, rr;   // entries of Sl and Sr
    if (cp) {
      l = s;
      rr = mkc(rp ? dg : 0.0, 0.0);
    } else {
      l = mkc(rp ? 0.0 : -dg, 0.0);
      rr = -s;
    }
    const cd t = sg * rr;
    cd a = l - t, b = l + t;
    if (flag[e / per]) {
      a = mkc(dg, 0.0);
      b = mkc(0.0, 0.0);
    }
    A[e] = a;
    B[e] = b;
  }
}

// Items whose Cayley matrix was singular (info < 0) get C = 0 so that the
// eigensolver sees finite entries.
__global__ __launch_bounds__(256) void zero_failed_kernel(const int* __restrict__ lu_info, cd* __restrict__ C,
                                                          long long per_item, long long total) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x)
    if (lu_info[e / per_item] != 0) C[e] = mkc(0.0, 0.0);
}

// w = s (c + 1)/(c - 1) in place; info = NaN code / LU status / eigensolver status.
__global__ __launch_bounds__(256) void cayley_finish_kernel(int n2, int batch, cd sg, const int* __restrict__ flag,
                                                            const int* __restrict__ lu_info, cd* __restrict__ w,
                                                            cd* __restrict__ v, int* __restrict__ info) {
  const long long total = (long long)batch * n2;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / n2;
    const int j = (int)(e % n2);
    const cd c = w[e];
    cd r = sg * divc(c + mkc(1.0, 0.0), c - mkc(1.0, 0.0));
    const bool bad = flag[b] != 0;
    if (bad || lu_info[b] != 0) r = mkc(NAN, NAN);
    w[e] = r;
    if (bad || lu_info[b] != 0) {
      cd* vb = v + b * (long long)n2 * n2;
      for (int i = 0; i < n2; ++i) vb[(long long)i * n2 + j] = mkc(NAN, NAN);
    }
    if (j == 0) info[b] = bad ? KHX_INFO_NAN_INPUT : (lu_info[b] != 0 ? lu_info[b] : info[b]);
  }
}

// det = sign * exp(sum log|u_ii|) over two LU factorisations (ipiv 0-based).
__global__ __launch_bounds__(64) void lu_det2_kernel(int n, const cd* __restrict__ L1, const int* __restrict__ p1,
                                                     const cd* __restrict__ L2, const int* __restrict__ p2,
                                                     cd* __restrict__ det) {
  const long long b = blockIdx.x;
  const long long sz = (long long)n * n;
  double lg = 0.0, ang = 0.0, swaps = 0.0, zero = 0.0;
  for (int i = threadIdx.x; i < n; i += 64) {
    const cd u1 = L1[b * sz + (long long)i * n + i], u2 = L2[b * sz + (long long)i * n + i];
    const double a1 = absc(u1), a2 = absc(u2);
    if (a1 == 0.0 || a2 == 0.0) zero += 1.0;
    lg += log(a1) + log(a2);
    ang += atan2(u1.y, u1.x) + atan2(u2.y, u2.x);
    swaps += (p1[b * n + i] != i ? 1.0 : 0.0) + (p2[b * n + i] != i ? 1.0 : 0.0);
  }
  lg = wave_sum(lg);
  ang = wave_sum(ang);
  swaps = wave_sum(swaps);
  zero = wave_sum(zero);
  if (threadIdx.x == 0) {
    cd r;
    if (zero > 0.0) {
      r = mkc(0.0, 0.0);
    } else {
      const double m = exp(lg) * ((((long long)swaps) & 1) ? -1.0 : 1.0);
      double sn, cs;
      sincos(ang, &sn, &cs);
      r = mkc(m * cs, m * sn);
    }
    det[b] = r;
  }
}

static unsigned grid_for(long long total) {
  long long g = (total + 255) / 256;
  return (unsigned)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

int scattering_splitlr(hipStream_t s, int ng, const cd* S, int batch, cd* Sl, cd* Sr) {
  KHX_REQUIRE(ng >= 1 && batch >= 0, "scattering_splitlr: bad sizes");
  const long long total = (long long)batch * 16 * ng * ng;
  if (total == 0) return 0;
  splitlr_kernel<<<grid_for(total), 256, 0, s>>>(2 * ng, S, Sl, Sr, total);
  KHX_LAUNCH_CHECK();
  return 0;
}

static size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t scattering_eig_ws_bytes(int ng, int batch) {
  const size_t n2 = 4 * (size_t)ng;
  return 2 * up256(n2 * n2 * sizeof(cd) * batch) + 2 * up256(sizeof(int) * batch) +
         up256(zeig_workspace_bytes((int)n2, batch));
}

int scattering_eig(hipStream_t s, int ng, const cd* S, int batch, cd sg, cd* w, cd* v, int* info, void* ws) {
  KHX_REQUIRE(ng >= 1 && batch >= 0 && 4ll * ng <= 8192, "scattering_eig: bad sizes");
  KHX_REQUIRE(fabs(absc(sg) - 1.0) < 1e-12, "scattering_eig: the shift must lie on the unit circle");
  if (batch == 0) return 0;
  const int n2 = 4 * ng;
  const long long per = (long long)n2 * n2, total = per * batch;
  char* p = static_cast<char*>(ws);
  cd* A = reinterpret_cast<cd*>(p);
  p += up256(sizeof(cd) * total);
  cd* C = reinterpret_cast<cd*>(p);
  p += up256(sizeof(cd) * total);
  int* flag = reinterpret_cast<int*>(p);
  p += up256(sizeof(int) * batch);
  int* lu_info = reinterpret_cast<int*>(p);
  p += up256(sizeof(int) * batch);
  void* eig_ws = p;
  fill_int_kernel<<<dim3((batch + 255) / 256), 256, 0, s>>>(flag, batch, 0);
  KHX_LAUNCH_CHECK();
  nan_flag_kernel<<<grid_for(total), 256, 0, s>>>(S, per, total, flag);
  KHX_LAUNCH_CHECK();
  cayley_kernel<<<grid_for(total), 256, 0, s>>>(2 * ng, S, sg, flag, A, C, total);
  KHX_LAUNCH_CHECK();
  KHX_TRY(zgesv(s, n2, n2, Mat{A, n2, per}, Mat{C, n2, per}, batch, lu_info));
  zero_failed_kernel<<<grid_for(total), 256, 0, s>>>(lu_info, C, per, total);
  KHX_LAUNCH_CHECK();
  KHX_TRY(zeig(s, n2, Mat{C, n2, per}, w, n2, Mat{v, n2, per}, batch, info, eig_ws, nullptr));
  cayley_finish_kernel<<<grid_for((long long)batch * n2), 256, 0, s>>>(n2, batch, sg, flag, lu_info, w, v, info);
  KHX_LAUNCH_CHECK();
  return 0;
}

size_t scattering_det_ws_bytes(int ng, int batch) {
  const size_t n = 2 * (size_t)ng;
  return 3 * up256(n * n * sizeof(cd) * batch) + 2 * up256(n * sizeof(int) * batch) + up256(sizeof(int) * batch);
}

int scattering_det(hipStream_t s, int ng, const cd* S, int batch, cd* det, int* info, void* ws) {
  KHX_REQUIRE(ng >= 1 && batch >= 0 && 2ll * ng <= 8192, "scattering_det: bad sizes");
  if (batch == 0) return 0;
  const int n = 2 * ng, n2 = 4 * ng;
  const long long per = (long long)n * n, sper = (long long)n2 * n2;
  char* p = static_cast<char*>(ws);
  cd* Amm = reinterpret_cast<cd*>(p);
  p += up256(sizeof(cd) * per * batch);
  cd* X = reinterpret_cast<cd*>(p);
  p += up256(sizeof(cd) * per * batch);
  cd* M11 = reinterpret_cast<cd*>(p);
  p += up256(sizeof(cd) * per * batch);
  int* piv1 = reinterpret_cast<int*>(p);
  p += up256(sizeof(int) * n * batch);
  int* piv2 = reinterpret_cast<int*>(p);
  p += up256(sizeof(int) * n * batch);
  int* info2 = reinterpret_cast<int*>(p);
  cd* Sm = const_cast<cd*>(S);
  const Mat Spp{Sm, n2, sper}, Spm{Sm + n, n2, sper}, Smp{Sm + (long long)n * n2, n2, sper},
      Smm{Sm + (long long)n * n2 + n, n2, sper};
  KHX_TRY(zcopy(s, Mat{Amm, n, per}, Smm, n, n, batch));
  KHX_TRY(zcopy(s, Mat{X, n, per}, Smp, n, n, batch));
  KHX_TRY(zcopy(s, Mat{M11, n, per}, Spp, n, n, batch));
  // Smm = P L U; X = Smm^-1 Smp; M11 = Spp - Spm X = P2 L2 U2
  KHX_TRY(zgetrf_solve(s, n, n, Mat{Amm, n, per}, Mat{X, n, per}, batch, piv1, info));
  KHX_TRY(zgemm(s, n, n, n, mkc(-1.0, 0.0), Spm, Mat{X, n, per}, mkc(1.0, 0.0), Mat{M11, n, per}, batch));
  KHX_TRY(zgetrf(s, n, Mat{M11, n, per}, batch, piv2, info2));
  lu_det2_kernel<<<dim3(batch), 64, 0, s>>>(n, Amm, piv1, M11, piv2, det);
  KHX_LAUNCH_CHECK();
  return 0;
}

}  // namespace khx
