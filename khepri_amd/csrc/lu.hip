// Batched LU factorisation with stored pivots and the matching solve.
//   zgetrf: A = P L U in place (unit L below the diagonal); ipiv[k] is the row
//           exchanged with row k at step k (0-based) -- LAPACK zgetrf layout.
//   zgetrs: X = A^-1 B from those factors (interchanges, then L and U sweeps).
// One workgroup per matrix; the pivot row and the multipliers of each step
// are staged in LDS.  Used where one factorisation serves several solves
// (the R/T chain applies D^-1 to vectors after building the next level).
#include "khx_internal.h"
#include "khx_prof.h"

#include <cstdlib>

namespace khx {

constexpr int LU_THREADS = 256;

__global__ __launch_bounds__(LU_THREADS) void zgetrf_kernel(int n, cd* A, int lda, long long sA, int* ipiv,
                                                            int* info) {
  extern __shared__ cd smem[];
  cd* prow = smem;      // n
  cd* lcol = smem + n;  // n
  __shared__ double red_v[LU_THREADS / 64];
  __shared__ int red_i[LU_THREADS / 64];
  __shared__ cd red_c[LU_THREADS / 64];
  __shared__ cd s_akk;
  const long long item = blockIdx.x;
  cd* a = A + item * sA;
  int* pv = ipiv + item * n;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  constexpr int NW = LU_THREADS / 64;
  int first_zero = 0;
  for (int k = 0; k < n; ++k) {
    double best = -1.0;
    int bi = n;
    cd bv = mkc(0.0, 0.0);
    for (int i = k + tid; i < n; i += LU_THREADS) {
      const cd v = a[(long long)i * lda + k];
      const double av = cabs1(v);
      if (av > best) {
        best = av;
        bi = i;
        bv = v;
      }
    }
    if (tid == 0) s_akk = a[(long long)k * lda + k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      const double ox = __shfl_xor(bv.x, o, 64), oy = __shfl_xor(bv.y, o, 64);
      if (ov > best || (ov == best && oi < bi)) {
        best = ov;
        bi = oi;
        bv = mkc(ox, oy);
      }
    }
    if (lane == 0) {
      red_v[wid] = best;
      red_i[wid] = bi;
      red_c[wid] = bv;
    }
    __syncthreads();
    best = red_v[0];
    bi = red_i[0];
    bv = red_c[0];
#pragma unroll
    for (int w = 1; w < NW; ++w)
      if (red_v[w] > best || (red_v[w] == best && red_i[w] < bi)) {
        best = red_v[w];
        bi = red_i[w];
        bv = red_c[w];
      }
    const cd akk = s_akk;
    int p = bi;
    cd piv = bv;
    if (p >= n) {
      p = k;
      piv = akk;
    }
    const bool zero = (piv.x == 0.0 && piv.y == 0.0);
    if (zero && first_zero == 0) first_zero = k + 1;
    const cd rp = zero ? mkc(0.0, 0.0) : recipc(piv);
    if (tid == 0) pv[k] = p;
    // full-row interchange (multipliers of earlier steps included)
    for (int j = tid; j < n; j += LU_THREADS) {
      const cd vk = a[(long long)k * lda + j], vp = a[(long long)p * lda + j];
      if (p != k) {
        a[(long long)k * lda + j] = vp;
        a[(long long)p * lda + j] = vk;
      }
      prow[j] = vp;
    }
    for (int i = k + 1 + tid; i < n; i += LU_THREADS) {
      const cd v = (i == p) ? akk : a[(long long)i * lda + k];
      lcol[i] = v * rp;
    }
    __syncthreads();
    for (int i = k + 1 + wid; i < n; i += NW) {
      const cd l = lcol[i];
      cd* ar = a + (long long)i * lda;
      if (lane == 0) ar[k] = l;
      for (int j = k + 1 + lane; j < n; j += 64) ar[j] = fmsc(ar[j], l, prow[j]);
    }
    __syncthreads();
  }
  if (tid == 0 && info) info[item] = first_zero;
}

__global__ __launch_bounds__(LU_THREADS) void zgetrs_kernel(int n, int m, const cd* A, int lda, long long sA,
                                                            const int* ipiv, cd* B, int ldb, long long sB) {
  extern __shared__ cd smem[];
  cd* xrow = smem;  // m
  const long long item = blockIdx.x;
  const cd* a = A + item * sA;
  const int* pv = ipiv + item * n;
  cd* bm = B + item * sB;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  constexpr int NW = LU_THREADS / 64;
  // interchanges in factorisation order: each thread owns whole columns of B
  for (int j = tid; j < m; j += LU_THREADS)
    for (int k = 0; k < n; ++k) {
      const int p = pv[k];
      if (p != k) {
        const cd t = bm[(long long)k * ldb + j];
        bm[(long long)k * ldb + j] = bm[(long long)p * ldb + j];
        bm[(long long)p * ldb + j] = t;
      }
    }
  __syncthreads();
  // L Y = P B (unit diagonal)
  for (int k = 0; k < n - 1; ++k) {
    for (int j = tid; j < m; j += LU_THREADS) xrow[j] = bm[(long long)k * ldb + j];
    __syncthreads();
    for (int i = k + 1 + wid; i < n; i += NW) {
      const cd l = a[(long long)i * lda + k];
      cd* br = bm + (long long)i * ldb;
      for (int j = lane; j < m; j += 64) br[j] = fmsc(br[j], l, xrow[j]);
    }
    __syncthreads();
  }
  // U X = Y
  for (int k = n - 1; k >= 0; --k) {
    const cd ukk = a[(long long)k * lda + k];
    for (int j = tid; j < m; j += LU_THREADS) {
      const cd x = divc(bm[(long long)k * ldb + j], ukk);
      bm[(long long)k * ldb + j] = x;
      xrow[j] = x;
    }
    __syncthreads();
    for (int i = wid; i < k; i += NW) {
      const cd u = a[(long long)i * lda + k];
      cd* br = bm + (long long)i * ldb;
      for (int j = lane; j < m; j += 64) br[j] = fmsc(br[j], u, xrow[j]);
    }
    __syncthreads();
  }
}

// Few right-hand sides (m <= ZGETRS_VEC_MAX): one wavefront per system.  The
// right-hand sides live in LDS; each substitution step is a row of L or U
// (contiguous, read once) dotted with the solved part, reduced across the
// wave -- no workgroup barrier on the sequential path.
constexpr int ZGETRS_VEC_MAX = 4;

__global__ __launch_bounds__(64) void zgetrs_vec_kernel(int n, int m, const cd* A, int lda, long long sA,
                                                        const int* ipiv, cd* B, int ldb, long long sB) {
  extern __shared__ cd xs[];  // n x m, row-major
  const long long item = blockIdx.x;
  const cd* a = A + item * sA;
  const int* pv = ipiv + item * n;
  cd* bm = B + item * sB;
  const int lane = threadIdx.x;
  for (int e = lane; e < n * m; e += 64) xs[e] = bm[(long long)(e / m) * ldb + e % m];
  __syncthreads();
  if (lane == 0)
    for (int k = 0; k < n; ++k) {
      const int p = pv[k];
      if (p != k)
        for (int j = 0; j < m; ++j) {
          const cd t = xs[k * m + j];
          xs[k * m + j] = xs[p * m + j];
          xs[p * m + j] = t;
        }
    }
  __syncthreads();
  // L y = P b (unit diagonal): y_i = b_i - sum_{k<i} L_ik y_k
  for (int i = 1; i < n; ++i) {
    const cd* row = a + (long long)i * lda;
    cd acc[ZGETRS_VEC_MAX];
#pragma unroll
    for (int j = 0; j < ZGETRS_VEC_MAX; ++j) acc[j] = mkc(0.0, 0.0);
    for (int k = lane; k < i; k += 64) {
      const cd l = row[k];
#pragma unroll
      for (int j = 0; j < ZGETRS_VEC_MAX; ++j)
        if (j < m) acc[j] = fmac(acc[j], l, xs[k * m + j]);
    }
#pragma unroll
    for (int j = 0; j < ZGETRS_VEC_MAX; ++j) {
      if (j >= m) break;
      const double re = wave_sum(acc[j].x), im = wave_sum(acc[j].y);
      if (lane == 0) xs[i * m + j] = xs[i * m + j] - mkc(re, im);
    }
    __syncthreads();
  }
  // U x = y: x_i = (y_i - sum_{k>i} U_ik x_k) / U_ii
  for (int i = n - 1; i >= 0; --i) {
    const cd* row = a + (long long)i * lda;
    cd acc[ZGETRS_VEC_MAX];
#pragma unroll
    for (int j = 0; j < ZGETRS_VEC_MAX; ++j) acc[j] = mkc(0.0, 0.0);
    for (int k = i + 1 + lane; k < n; k += 64) {
      const cd u = row[k];
#pragma unroll
      for (int j = 0; j < ZGETRS_VEC_MAX; ++j)
        if (j < m) acc[j] = fmac(acc[j], u, xs[k * m + j]);
    }
    const cd uii = row[i];
#pragma unroll
    for (int j = 0; j < ZGETRS_VEC_MAX; ++j) {
      if (j >= m) break;
      const double re = wave_sum(acc[j].x), im = wave_sum(acc[j].y);
      if (lane == 0) xs[i * m + j] = divc(xs[i * m + j] - mkc(re, im), uii);
    }
    __syncthreads();
  }
  for (int e = lane; e < n * m; e += 64) bm[(long long)(e / m) * ldb + e % m] = xs[e];
}

int zgetrf(hipStream_t s, int n, Mat A, int batch, int* ipiv, int* info) {
  if (batch <= 0 || n <= 0) return 0;
  const int path = lu_path(n);
  if (path == 2) return zlu_multi(s, n, 0, A, Mat{nullptr, 0, 0}, batch, ipiv, info, true);
  if (path == 1) return zlu_blocked(s, n, 0, A, Mat{nullptr, 0, 0}, batch, ipiv, info, true);
  const size_t lds = (size_t)2 * n * sizeof(cd);
  KHX_REQUIRE(lds <= 64 * 1024, "zgetrf: n too large for the LDS staging buffers");
  ProfScope ps("zgetrf_kernel", s);
  zgetrf_kernel<<<dim3(batch), dim3(LU_THREADS), lds, s>>>(n, A.p, A.ld, A.stride, ipiv, info);
  KHX_LAUNCH_CHECK();
  return 0;
}

// Factor and solve in one pass: A -> P L U (zgetrf layout, ipiv), B -> A^-1 B.
int zgetrf_solve(hipStream_t s, int n, int m, Mat A, Mat B, int batch, int* ipiv, int* info) {
  if (batch <= 0 || n <= 0) return 0;
  const int path = lu_path(n);
  if (path == 2) return zlu_multi(s, n, m, A, B, batch, ipiv, info, true);
  if (path == 1) return zlu_blocked(s, n, m, A, B, batch, ipiv, info, true);
  KHX_TRY(zgetrf(s, n, A, batch, ipiv, info));
  return zgetrs(s, n, m, A, ipiv, B, batch);
}

int zgetrs(hipStream_t s, int n, int m, Mat A, const int* ipiv, Mat B, int batch) {
  if (batch <= 0 || n <= 0 || m <= 0) return 0;
  // large systems: the sequential row-by-row solves below leave the GPU idle
  // (one wavefront per system); split the triangles into GEMMs instead
  if (n >= 200 && batch <= 65535 && (A.stride != 0 || batch == 1) && (B.stride != 0 || batch == 1))
    return zlu_solve_factored(s, n, m, A, ipiv, B, batch);
  if (m <= ZGETRS_VEC_MAX && (size_t)n * m * sizeof(cd) <= 64 * 1024) {
    ProfScope ps("zgetrs_kernel", s);
    zgetrs_vec_kernel<<<dim3(batch), dim3(64), (size_t)n * m * sizeof(cd), s>>>(n, m, A.p, A.ld, A.stride, ipiv,
                                                                              B.p, B.ld, B.stride);
    KHX_LAUNCH_CHECK();
    return 0;
  }
  const size_t lds = (size_t)m * sizeof(cd);
  KHX_REQUIRE(lds <= 64 * 1024, "zgetrs: too many right-hand sides for the LDS row buffer");
  ProfScope ps("zgetrs_kernel", s);
  zgetrs_kernel<<<dim3(batch), dim3(LU_THREADS), lds, s>>>(n, m, A.p, A.ld, A.stride, ipiv, B.p, B.ld, B.stride);
  KHX_LAUNCH_CHECK();
  return 0;
}

}  // namespace khx
