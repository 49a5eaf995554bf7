// Blocked LU solve / factorisation, one workgroup per system.
//
// Right-looking LU with partial pivoting in panels of NB columns:
//   1. the panel A[j0:n, j0:j0+NB] is factorised in LDS (pivot search, row
//      interchange, scaling and rank-1 updates never leave LDS);
//   2. its interchanges are applied to the other columns (and to B);
//   3. U12 = L11^-1 A12 and Y1 = L11^-1 B1 (one thread per column);
//   4. the trailing block [A22 | B2] -= L21 [U12 | Y1] as an LDS-staged
//      mini-GEMM (L21 from the panel, [U12 | Y1] in 64-column tiles, 4x4
//      complex register tiles per thread).
// With right-hand sides, the back substitution U X = Y is blocked the same
// way (NB x NB diagonal solves, then mini-GEMM updates of the rows above).
// Every element of the trailing matrix is read and written once per panel
// instead of once per column.
#include <mutex>
#include <utility>

#include "khx_internal.h"
#include "khx_prof.h"

namespace khx {

constexpr int LB_THREADS = 256;
constexpr int LB_NB = 16;  // panel width
constexpr int LB_PS = LB_NB + 1;  // LDS row stride of a panel (padding: one row per thread hits distinct banks)
constexpr int LB_TW = 64;  // column tile of the mini-GEMM
constexpr int LB_SG = 4;   // interchanges per independent group in the panel kernel (the code spells out 4)

// C[r][c] -= sum_k L[r - r0][k] * U[k][c - cbase] for rows r in [r0, r1) and
// tile columns [0, tw); L rows live in LDS with stride LB_PS, the U tile in
// LDS with stride LB_TW.  colptr(t) gives the global address of (row 0,
// tile column t); rows are ld apart.
template <typename ColPtr>
__device__ __forceinline__ void mini_gemm_sub(const cd* Ls, int lrow0, const cd* Us, int nb, int r0, int r1, int tw,
                                              ColPtr colptr, long long ld) {
  const int tid = threadIdx.x;
  const int cg = tid & 15;  // 16 column groups of 4
  const int rg = tid >> 4;  // 16 row groups of 4
  const int cbase = cg * 4;
  for (int rb = r0 + rg * 4; rb < r1; rb += 64) {
    cd acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = mkc(0.0, 0.0);
    for (int kk = 0; kk < nb; ++kk) {
      cd lv[4], uv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) lv[i] = (rb + i < r1) ? Ls[(rb + i - lrow0) * LB_PS + kk] : mkc(0.0, 0.0);
#pragma unroll
      for (int j = 0; j < 4; ++j) uv[j] = Us[kk * LB_TW + cbase + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmac(acc[i][j], lv[i], uv[j]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (cbase + j >= tw) continue;
      cd* col = colptr(cbase + j);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (rb + i >= r1) continue;
        cd* p = col + (long long)(rb + i) * ld;
        *p = *p - acc[i][j];
      }
    }
  }
}

template <int TH = LB_THREADS>
__device__ __forceinline__ void argmax_abs1(double& best, int& bi, cd& bv, double* rv, int* ri, cd* rc) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
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
    rv[wid] = best;
    ri[wid] = bi;
    rc[wid] = bv;
  }
  __syncthreads();
  best = rv[0];
  bi = ri[0];
  bv = rc[0];
#pragma unroll
  for (int w = 1; w < TH / 64; ++w)
    if (rv[w] > best || (rv[w] == best && ri[w] < bi)) {
      best = rv[w];
      bi = ri[w];
      bv = rc[w];
    }
}

// keepL: LAPACK zgetrf layout (interchanges applied to the L part too, ipiv
// stored).  m > 0: B (n x m) is transformed along and finally solved.
__global__ __launch_bounds__(LB_THREADS) void zlu_blocked_kernel(int n, int m, cd* A, int lda, long long sA, cd* B,
                                                                 int ldb, long long sB, int* ipiv, int* info,
                                                                 int keepL) {
  extern __shared__ cd smem[];
  cd* pan = smem;                // n x LB_PS
  cd* ut = pan + n * LB_PS;      // LB_NB x LB_TW
  __shared__ double rv[LB_THREADS / 64];
  __shared__ int ri[LB_THREADS / 64];
  __shared__ cd rc[LB_THREADS / 64];
  __shared__ int perm[LB_NB];
  const long long item = blockIdx.x;
  cd* a = A + item * sA;
  cd* bm = m > 0 ? B + item * sB : nullptr;
  const int tid = threadIdx.x;
  int first_zero = 0;

  for (int j0 = 0; j0 < n; j0 += LB_NB) {
    const int nb = min(LB_NB, n - j0);
    const int pr = n - j0;
    for (int e = tid; e < pr * LB_NB; e += LB_THREADS) {
      const int r = e / LB_NB, c = e % LB_NB;
      pan[r * LB_PS + c] = (c < nb) ? a[(long long)(j0 + r) * lda + j0 + c] : mkc(0.0, 0.0);
    }
    __syncthreads();
    // ---- factor the panel in LDS
    for (int jj = 0; jj < nb; ++jj) {
      double best = -1.0;
      int bi = pr;
      cd bv = mkc(0.0, 0.0);
      for (int r = jj + tid; r < pr; r += LB_THREADS) {
        const cd v = pan[r * LB_PS + jj];
        const double av = cabs1(v);
        if (av > best) {
          best = av;
          bi = r;
          bv = v;
        }
      }
      argmax_abs1(best, bi, bv, rv, ri, rc);
      int p = bi;
      cd piv = bv;
      if (p >= pr) {  // NaN column: keep the diagonal
        p = jj;
        piv = pan[jj * LB_PS + jj];
      }
      const bool zero = (piv.x == 0.0 && piv.y == 0.0);
      if (zero && first_zero == 0) first_zero = j0 + jj + 1;
      const cd rp = zero ? mkc(0.0, 0.0) : recipc(piv);
      if (tid == 0) {
        perm[jj] = p;
        if (ipiv) ipiv[item * n + j0 + jj] = j0 + p;
      }
      __syncthreads();  // every thread has read the pivot column before the swap
      if (p != jj && tid < LB_NB) {
        const cd t = pan[jj * LB_PS + tid];
        pan[jj * LB_PS + tid] = pan[p * LB_PS + tid];
        pan[p * LB_PS + tid] = t;
      }
      __syncthreads();
      for (int r = jj + 1 + tid; r < pr; r += LB_THREADS) {
        cd* row = pan + r * LB_PS;
        const cd l = row[jj] * rp;
        row[jj] = l;
        for (int c = jj + 1; c < nb; ++c) row[c] = fmsc(row[c], l, pan[jj * LB_PS + c]);
      }
      __syncthreads();
    }
    for (int e = tid; e < pr * LB_NB; e += LB_THREADS) {
      const int r = e / LB_NB, c = e % LB_NB;
      if (c < nb) a[(long long)(j0 + r) * lda + j0 + c] = pan[r * LB_PS + c];
    }
    // ---- interchanges on the other columns (one thread per column)
    const int nleft = keepL ? j0 : 0;
    const int nright = n - j0 - nb;
    for (int e = tid; e < nleft + nright + m; e += LB_THREADS) {
      cd* col;
      long long ld;
      if (e < nleft) {
        col = a + e;
        ld = lda;
      } else if (e < nleft + nright) {
        col = a + j0 + nb + (e - nleft);
        ld = lda;
      } else {
        col = bm + (e - nleft - nright);
        ld = ldb;
      }
      for (int jj = 0; jj < nb; ++jj) {
        const int p = perm[jj];
        if (p != jj) {
          cd* x = col + (long long)(j0 + jj) * ld;
          cd* y = col + (long long)(j0 + p) * ld;
          const cd t = *x;
          *x = *y;
          *y = t;
        }
      }
    }
    __syncthreads();
    // ---- U12 = L11^-1 A12, Y1 = L11^-1 B1 (unit lower, one thread per column)
    for (int e = tid; e < nright + m; e += LB_THREADS) {
      cd* col = (e < nright) ? a + j0 + nb + e : bm + (e - nright);
      const long long ld = (e < nright) ? lda : ldb;
      cd x[LB_NB];
#pragma unroll
      for (int i = 0; i < LB_NB; ++i) x[i] = (i < nb) ? col[(long long)(j0 + i) * ld] : mkc(0.0, 0.0);
#pragma unroll
      for (int i = 1; i < LB_NB; ++i)
#pragma unroll
        for (int k = 0; k < i; ++k) x[i] = fmsc(x[i], pan[i * LB_PS + k], x[k]);
#pragma unroll
      for (int i = 0; i < LB_NB; ++i)
        if (i < nb) col[(long long)(j0 + i) * ld] = x[i];
    }
    __syncthreads();
    // ---- trailing update [A22 | B2] -= L21 [U12 | Y1]
    if (j0 + nb < n) {
      const int tot = nright + m;
      for (int c0 = 0; c0 < tot; c0 += LB_TW) {
        const int tw = min(LB_TW, tot - c0);
        for (int e = tid; e < LB_NB * LB_TW; e += LB_THREADS) {
          const int kk = e / LB_TW, t = e % LB_TW;
          cd v = mkc(0.0, 0.0);
          if (kk < nb && t < tw) {
            const int c = c0 + t;
            v = (c < nright) ? a[(long long)(j0 + kk) * lda + j0 + nb + c]
                             : bm[(long long)(j0 + kk) * ldb + (c - nright)];
          }
          ut[e] = v;
        }
        __syncthreads();
        // A and B columns have different leading dimensions: the tile is
        // split at the A/B boundary.
        const int ta = max(0, min(tw, nright - c0));
        if (ta > 0)
          mini_gemm_sub(pan, 0, ut, nb, nb, pr, ta,
                        [&](int t) { return a + (long long)j0 * lda + j0 + nb + c0 + t; }, lda);
        if (ta < tw) {
          // shift the tile so that B columns start at tile column 0
          __syncthreads();
          for (int e = tid; e < LB_NB * LB_TW; e += LB_THREADS) {
            const int kk = e / LB_TW, t = e % LB_TW;
            cd v = mkc(0.0, 0.0);
            if (kk < nb && ta + t < tw) v = bm[(long long)(j0 + kk) * ldb + (c0 + ta + t - nright)];
            ut[e] = v;
          }
          __syncthreads();
          mini_gemm_sub(pan, 0, ut, nb, nb, pr, tw - ta,
                        [&](int t) { return bm + (long long)j0 * ldb + (c0 + ta + t - nright); }, ldb);
        }
        __syncthreads();
      }
    }
  }
  if (tid == 0 && info) info[item] = first_zero;
  if (m <= 0) return;

  // ---- back substitution U X = Y, NB rows at a time from the bottom
  const int nblk = (n + LB_NB - 1) / LB_NB;
  for (int blk = nblk - 1; blk >= 0; --blk) {
    const int j0 = blk * LB_NB;
    const int nb = min(LB_NB, n - j0);
    // U[0:j0+nb, j0:j0+nb] -> pan (rows 0..j0+nb-1)
    for (int e = tid; e < (j0 + nb) * LB_NB; e += LB_THREADS) {
      const int r = e / LB_NB, c = e % LB_NB;
      pan[r * LB_PS + c] = (c < nb) ? a[(long long)r * lda + j0 + c] : mkc(0.0, 0.0);
    }
    __syncthreads();
    // X block = U11^-1 Y block (one thread per column of B)
    for (int c = tid; c < m; c += LB_THREADS) {
      cd* col = bm + c;
      cd x[LB_NB];
#pragma unroll
      for (int i = 0; i < LB_NB; ++i) x[i] = (i < nb) ? col[(long long)(j0 + i) * ldb] : mkc(0.0, 0.0);
#pragma unroll
      for (int i = LB_NB - 1; i >= 0; --i) {
        if (i >= nb) continue;
#pragma unroll
        for (int k = i + 1; k < LB_NB; ++k)
          if (k < nb) x[i] = fmsc(x[i], pan[(j0 + i) * LB_PS + k], x[k]);
        x[i] = divc(x[i], pan[(j0 + i) * LB_PS + i]);
      }
#pragma unroll
      for (int i = 0; i < LB_NB; ++i)
        if (i < nb) col[(long long)(j0 + i) * ldb] = x[i];
    }
    __syncthreads();
    // Y[0:j0] -= U[0:j0, j0:j0+nb] X block
    if (j0 > 0) {
      for (int c0 = 0; c0 < m; c0 += LB_TW) {
        const int tw = min(LB_TW, m - c0);
        for (int e = tid; e < LB_NB * LB_TW; e += LB_THREADS) {
          const int kk = e / LB_TW, t = e % LB_TW;
          ut[e] = (kk < nb && t < tw) ? bm[(long long)(j0 + kk) * ldb + c0 + t] : mkc(0.0, 0.0);
        }
        __syncthreads();
        mini_gemm_sub(pan, 0, ut, nb, 0, j0, tw, [&](int t) { return bm + c0 + t; }, ldb);
        __syncthreads();
      }
    }
  }
}

// After a panel is factored: its status word, then its interchanges on the
// other columns (and B) and the triangular solve with L11 (pan: the panel's
// first rows, unit lower, row stride pld), one thread per column.
template <int NB, int TH = LB_THREADS>
__device__ __forceinline__ void panel_apply(int n, int m, int j0, int nb, cd* a, int lda, cd* bm, int ldb,
                                            const cd* pan, long long pld, const int* perm, int* info,
                                            long long item, int first_zero, int keepL, int cend) {
  const int tid = threadIdx.x;
  if (tid == 0 && info) {
    if (j0 == 0)
      info[item] = first_zero;
    else if (info[item] == 0)
      info[item] = first_zero;
  }
  // The interchanges go in groups of LB_SG: when the rows a group touches
  // are distinct (no pivot row is another row of the group), its swaps are
  // independent and each column loads them all before storing, instead of
  // one dependent load / store round per interchange.
  __shared__ int gfree[NB / LB_SG];
  if (tid < NB / LB_SG) {
    bool ok = true;
    const int g0 = tid * LB_SG;
    for (int i = g0; i < g0 + LB_SG && i < nb; ++i) {
      const int pi = perm[i];
      for (int k = g0; k < g0 + LB_SG && k < nb; ++k)
        if (k != i && (pi == k || (pi != i && perm[k] == pi))) ok = false;
    }
    gfree[tid] = ok ? 1 : 0;
  }
  __syncthreads();
  // interchanges on the other columns and the triangular solve with L11,
  // one thread per column
  const int nleft = (keepL & 1) ? j0 : 0;
  const int nright = n - j0 - nb;
  for (int e = tid; e < nleft + nright + m; e += TH) {
    cd* col;
    long long ld;
    if (e < nleft) {
      col = a + e;
      ld = lda;
    } else if (e < nleft + nright) {
      col = a + j0 + nb + (e - nleft);
      ld = lda;
    } else {
      col = bm + (e - nleft - nright);
      ld = ldb;
    }
    cd* c0p = col + (long long)j0 * ld;
    for (int g0 = 0; g0 < nb; g0 += LB_SG) {
      if (gfree[g0 / LB_SG]) {
        const cd z = mkc(0.0, 0.0);
        cd xa0 = z, xa1 = z, xa2 = z, xa3 = z, xb0 = z, xb1 = z, xb2 = z, xb3 = z;
        const int p0 = g0 < nb ? perm[g0] : g0, p1 = g0 + 1 < nb ? perm[g0 + 1] : g0 + 1;
        const int p2 = g0 + 2 < nb ? perm[g0 + 2] : g0 + 2, p3 = g0 + 3 < nb ? perm[g0 + 3] : g0 + 3;
        if (p0 != g0) { xa0 = c0p[(long long)g0 * ld]; xb0 = c0p[(long long)p0 * ld]; }
        if (p1 != g0 + 1) { xa1 = c0p[(long long)(g0 + 1) * ld]; xb1 = c0p[(long long)p1 * ld]; }
        if (p2 != g0 + 2) { xa2 = c0p[(long long)(g0 + 2) * ld]; xb2 = c0p[(long long)p2 * ld]; }
        if (p3 != g0 + 3) { xa3 = c0p[(long long)(g0 + 3) * ld]; xb3 = c0p[(long long)p3 * ld]; }
        if (p0 != g0) { c0p[(long long)g0 * ld] = xb0; c0p[(long long)p0 * ld] = xa0; }
        if (p1 != g0 + 1) { c0p[(long long)(g0 + 1) * ld] = xb1; c0p[(long long)p1 * ld] = xa1; }
        if (p2 != g0 + 2) { c0p[(long long)(g0 + 2) * ld] = xb2; c0p[(long long)p2 * ld] = xa2; }
        if (p3 != g0 + 3) { c0p[(long long)(g0 + 3) * ld] = xb3; c0p[(long long)p3 * ld] = xa3; }
      } else {
        for (int jj = g0; jj < g0 + LB_SG && jj < nb; ++jj) {
          const int p = perm[jj];
          if (p != jj) {
            cd* x = c0p + (long long)jj * ld;
            cd* y = c0p + (long long)p * ld;
            const cd t = *x;
            *x = *y;
            *y = t;
          }
        }
      }
    }
    const bool trsm = (e >= nleft && e < nleft + nright) ? (j0 + nb + (e - nleft) < cend) : (e >= nleft && !(keepL & 2));
    if (trsm) {
      cd x[NB];
#pragma unroll
      for (int i = 0; i < NB; ++i) x[i] = (i < nb) ? col[(long long)(j0 + i) * ld] : mkc(0.0, 0.0);
#pragma unroll
      for (int i = 1; i < NB; ++i)
        if (i < nb) {  // the last panel may be narrower: rows past it are outside the matrix
#pragma unroll
          for (int k = 0; k < i; ++k) x[i] = fmsc(x[i], pan[i * pld + k], x[k]);
        }
#pragma unroll
      for (int i = 0; i < NB; ++i)
        if (i < nb) col[(long long)(j0 + i) * ld] = x[i];
    }
  }
}

// ------------------------------------------------------------------ multi-launch variant
// The same right-looking LU, one panel per launch: zlu_panel_kernel factors
// the panel A[j0:n, j0:j0+16] in LDS, applies its interchanges to the rest of
// the rows (and to B), and solves U12 = L11^-1 A12, Y1 = L11^-1 B1; the
// trailing update [A22 | B2] -= L21 [U12 | Y1] is a batched zgemm launch, so
// the bulk of the flops runs in 64 x 64 GEMM tiles spread over the whole GPU
// instead of inside the one workgroup that owns the matrix.
// GP: the panel stays in global memory (systems too large for an LDS panel).
// keepL bit 0: apply the interchanges to the factored columns left of the
// panel too (zgetrf layout); bit 1: only interchange B's rows (its triangular
// solves run after the factorization).  Columns right of the panel are all
// interchanged; only those before cend are also solved with L11.
template <int NB, bool GP>
__global__ __launch_bounds__(LB_THREADS) void zlu_panel_kernel(int n, int m, int j0, cd* A, int lda, long long sA,
                                                               cd* B, int ldb, long long sB, int* ipiv, int* info,
                                                               int keepL, int cend) {
  extern __shared__ cd smem[];
  __shared__ double rv[LB_THREADS / 64];
  __shared__ int ri[LB_THREADS / 64];
  __shared__ cd rc[LB_THREADS / 64];
  __shared__ int perm[NB];
  const long long item = blockIdx.x;
  cd* a = A + item * sA;
  cd* bm = m > 0 ? B + item * sB : nullptr;
  const int tid = threadIdx.x;
  int first_zero = 0;
  const int nb = min(NB, n - j0);
  const int pr = n - j0;
  // panel element (r, c): LDS copy (row stride NB + 1) or the matrix itself
  cd* pan = GP ? a + (long long)j0 * lda + j0 : smem;
  // LDS rows padded by one element: a thread per row then hits distinct banks
  const long long pld = GP ? lda : NB + 1;
  if (!GP) {
    for (int e = tid; e < pr * NB; e += LB_THREADS) {
      const int r = e / NB, c = e % NB;
      pan[r * pld + c] = (c < nb) ? a[(long long)(j0 + r) * lda + j0 + c] : mkc(0.0, 0.0);
    }
  }
  __syncthreads();
  for (int jj = 0; jj < nb; ++jj) {
    double best = -1.0;
    int bi = pr;
    cd bv = mkc(0.0, 0.0);
    for (int r = jj + tid; r < pr; r += LB_THREADS) {
      const cd v = pan[r * pld + jj];
      const double av = cabs1(v);
      if (av > best) {
        best = av;
        bi = r;
        bv = v;
      }
    }
    argmax_abs1(best, bi, bv, rv, ri, rc);
    int p = bi;
    cd piv = bv;
    if (p >= pr) {  // NaN column: keep the diagonal
      p = jj;
      piv = pan[jj * pld + jj];
    }
    const bool zero = (piv.x == 0.0 && piv.y == 0.0);
    if (zero && first_zero == 0) first_zero = j0 + jj + 1;
    const cd rp = zero ? mkc(0.0, 0.0) : recipc(piv);
    if (tid == 0) {
      perm[jj] = p;
      if (ipiv) ipiv[item * n + j0 + jj] = j0 + p;
    }
    __syncthreads();  // every thread has read the pivot column before the swap
    if (p != jj && tid < nb) {
      const cd t = pan[jj * pld + tid];
      pan[jj * pld + tid] = pan[p * pld + tid];
      pan[p * pld + tid] = t;
    }
    __syncthreads();
    for (int r = jj + 1 + tid; r < pr; r += LB_THREADS) {
      cd* row = pan + r * pld;
      const cd l = row[jj] * rp;
      row[jj] = l;
      for (int c = jj + 1; c < nb; ++c) row[c] = fmsc(row[c], l, pan[jj * pld + c]);
    }
    __syncthreads();
  }
  if (!GP) {
    for (int e = tid; e < pr * NB; e += LB_THREADS) {
      const int r = e / NB, c = e % NB;
      if (c < nb) a[(long long)(j0 + r) * lda + j0 + c] = pan[r * pld + c];
    }
  }
  panel_apply<NB>(n, m, j0, nb, a, lda, bm, ldb, pan, pld, perm, info, item, first_zero, keepL, cend);
}

// Pivot step JJ of the register panel (a template parameter, so that every
// index into the row registers is a constant).
template <int NB, int RPT, int JJ, int TH>
__device__ __forceinline__ void panel_step(cd (&v)[RPT][NB], int nb, int pr, int j0, int n, long long item, int* ipiv,
                                           int* perm, cd* prow, cd* jrow, double* rv, int* ri, cd* rc,
                                           int& first_zero) {
  if (JJ >= nb) return;  // uniform: the last panel may be narrower
  const int tid = threadIdx.x;
  double best = -1.0;
  int bi = pr;
  cd bv = mkc(0.0, 0.0);
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    const int r = tid + k * TH;
    if (r >= JJ && r < pr) {
      const double av = cabs1(v[k][JJ]);
      if (av > best) {
        best = av;
        bi = r;
        bv = v[k][JJ];
      }
    }
  }
  argmax_abs1<TH>(best, bi, bv, rv, ri, rc);
  const int p = bi < pr ? bi : JJ;  // NaN column: keep the diagonal
  // rows p (the new pivot row) and JJ trade places through LDS
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    const int r = tid + k * TH;
    if (r == p) {
#pragma unroll
      for (int c = 0; c < NB; ++c) prow[c] = v[k][c];
    } else if (r == JJ) {
#pragma unroll
      for (int c = 0; c < NB; ++c) jrow[c] = v[k][c];
    }
  }
  if (tid == 0) {
    perm[JJ] = p;
    if (ipiv) ipiv[item * n + j0 + JJ] = j0 + p;
  }
  __syncthreads();
  const cd piv = prow[JJ];
  const bool zero = (piv.x == 0.0 && piv.y == 0.0);
  if (zero && first_zero == 0) first_zero = j0 + JJ + 1;
  const cd rp = zero ? mkc(0.0, 0.0) : recipc(piv);
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    const int r = tid + k * TH;
    if (r == JJ || r == p) {  // one set of loads: row JJ takes the pivot row, row p the old row JJ
      const cd* src = (r == JJ) ? prow : jrow;
#pragma unroll
      for (int c = 0; c < NB; ++c) v[k][c] = src[c];
    }
    if (r > JJ && r < pr) {
      const cd l = v[k][JJ] * rp;
      v[k][JJ] = l;
#pragma unroll
      for (int c = JJ + 1; c < NB; ++c) v[k][c] = fmsc(v[k][c], l, prow[c]);
    }
  }
}

template <int NB, int RPT, int TH, int... J>
__device__ __forceinline__ void panel_steps(std::integer_sequence<int, J...>, cd (&v)[RPT][NB], int nb, int pr, int j0,
                                            int n, long long item, int* ipiv, int* perm, cd* prow, cd* jrow,
                                            double* rv, int* ri, cd* rc, int& first_zero) {
  (panel_step<NB, RPT, J, TH>(v, nb, pr, j0, n, item, ipiv, perm, prow, jrow, rv, ri, rc, first_zero), ...);
}

// The same panel step with the panel in registers: thread t holds the
// panel rows t, t + 256, .. (RPT of them, pr <= 256 RPT) and every pivot
// step runs on registers, two barriers per column (the pivot search's
// cross-wave reduction, then the exchange of rows jj and p through LDS,
// which also broadcasts the pivot row).  Pivot choice, arithmetic and
// output are those of zlu_panel_kernel: the largest |Re| + |Im| below the
// diagonal, the first such row on ties, NaN columns keep the diagonal.
template <int NB, int RPT, int TH = LB_THREADS>
__global__ __launch_bounds__(TH, TH == 256 ? ((RPT == 1 && NB == 16) ? 4 : 2) : 1) void zlu_panel_reg_kernel(int n, int m, int j0, cd* A, int lda, long long sA,
                                                                   cd* B, int ldb, long long sB, int* ipiv, int* info,
                                                                   int keepL, int cend) {
  __shared__ double rv[TH / 64];
  __shared__ int ri[TH / 64];
  __shared__ cd rc[TH / 64];
  __shared__ int perm[NB];
  __shared__ cd prow[NB], jrow[NB];
  // L11 for the column solves, in dynamic LDS (NB (NB + 1) elements): with a
  // static array the compiler hoists all its loads out of the column loop
  extern __shared__ cd l11[];
  const long long item = blockIdx.x;
  cd* a = A + item * sA;
  cd* bm = m > 0 ? B + item * sB : nullptr;
  const int tid = threadIdx.x;
  int first_zero = 0;
  const int nb = min(NB, n - j0);
  const int pr = n - j0;
  cd v[RPT][NB];
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    const int r = tid + k * TH;
    const cd* src = a + (long long)(j0 + r) * lda + j0;
#pragma unroll
    for (int c = 0; c < NB; ++c) v[k][c] = (r < pr && c < nb) ? src[c] : mkc(0.0, 0.0);
  }
  panel_steps<NB, RPT, TH>(std::make_integer_sequence<int, NB>{}, v, nb, pr, j0, n, item, ipiv, perm, prow, jrow, rv, ri,
                        rc, first_zero);
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    const int r = tid + k * TH;
    if (r < pr) {
      cd* dst = a + (long long)(j0 + r) * lda + j0;
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < nb) dst[c] = v[k][c];
      if (r < NB) {
#pragma unroll
        for (int c = 0; c < NB; ++c) l11[r * (NB + 1) + c] = v[k][c];
      }
    }
  }
  __syncthreads();
  panel_apply<NB, TH>(n, m, j0, nb, a, lda, bm, ldb, l11, NB + 1, perm, info, item, first_zero, keepL, cend);
}

// X block = U11^-1 Y block for the rows j0 .. j0+nb of B (one thread per column)
template <int NB>
__global__ __launch_bounds__(LB_THREADS) void zlu_upper_block_kernel(int n, int m, int j0, const cd* A, int lda,
                                                                     long long sA, cd* B, int ldb, long long sB) {
  __shared__ cd u[NB * NB];
  const long long item = blockIdx.y;
  const cd* a = A + item * sA;
  cd* bm = B + item * sB;
  const int nb = min(NB, n - j0);
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) {
    const int r = e / NB, c = e % NB;
    u[e] = (r < nb && c < nb) ? a[(long long)(j0 + r) * lda + j0 + c] : mkc(r == c ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  cd* col = bm + c;
  cd x[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) x[i] = (i < nb) ? col[(long long)(j0 + i) * ldb] : mkc(0.0, 0.0);
#pragma unroll
  for (int i = NB - 1; i >= 0; --i) {
#pragma unroll
    for (int k = i + 1; k < NB; ++k) x[i] = fmsc(x[i], u[i * NB + k], x[k]);
    x[i] = divc(x[i], u[i * NB + i]);
  }
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < nb) col[(long long)(j0 + i) * ldb] = x[i];
}

// X block = L11^-1 Y block (unit lower) for the rows j0 .. j0+nb of B (one thread per column)
template <int NB>
__global__ __launch_bounds__(LB_THREADS) void zlu_lower_block_kernel(int n, int m, int j0, const cd* A, int lda,
                                                                     long long sA, cd* B, int ldb, long long sB) {
  __shared__ cd lw[NB * NB];
  const long long item = blockIdx.y;
  const cd* a = A + item * sA;
  cd* bm = B + item * sB;
  const int nb = min(NB, n - j0);
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) {
    const int r = e / NB, c = e % NB;
    lw[e] = (r < nb && c < r) ? a[(long long)(j0 + r) * lda + j0 + c] : mkc(0.0, 0.0);
  }
  __syncthreads();
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  cd* col = bm + c;
  cd x[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) x[i] = (i < nb) ? col[(long long)(j0 + i) * ldb] : mkc(0.0, 0.0);
#pragma unroll
  for (int i = 1; i < NB; ++i)
#pragma unroll
    for (int k = 0; k < i; ++k) x[i] = fmsc(x[i], lw[i * NB + k], x[k]);
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < nb) col[(long long)(j0 + i) * ldb] = x[i];
}

// Triangular solves with the factors of the multi-launch LU, split in halves
// so that most of the work is one GEMM with a deep K (rows r0 .. r1 of B):
//   L: X1 = L11^-1 B1;  B2 -= L21 X1;  X2 = L22^-1 B2
//   U: X2 = U22^-1 B2;  B1 -= U12 X2;  X1 = U11^-1 B1
// Ranges of at most 2 NB rows use the NB-row block kernels directly.
// GEMM of the triangular-solve updates: FP64 matrix cores, or FP32 ones in
// the mixed-precision solve (zgesv_refined)
template <bool F32>
static int trsm_gemm(hipStream_t s, int M, int N, int K, cd alpha, Mat A, Mat B, cd beta, Mat C, int batch) {
  if constexpr (F32) return zgemm_f32(s, M, N, K, alpha, A, B, beta, C, batch);
  return zgemm(s, M, N, K, alpha, A, B, beta, C, batch);
}

template <int NB, bool F32 = false>
static int trsm_lower_rec(hipStream_t s, int n, int m, Mat A, Mat B, int batch, int r0, int r1) {
  const cd one = mkc(1.0, 0.0), mone = mkc(-1.0, 0.0);
  if (r1 - r0 <= 2 * NB) {
    for (int j0 = r0; j0 < r1; j0 += NB) {
      const int nb = min(NB, r1 - j0);
      zlu_lower_block_kernel<NB><<<dim3((m + LB_THREADS - 1) / LB_THREADS, batch), dim3(LB_THREADS), 0, s>>>(
          min(n, j0 + nb), m, j0, A.p, A.ld, A.stride, B.p, B.ld, B.stride);
      KHX_LAUNCH_CHECK();
      if (j0 + nb < r1)
        KHX_TRY(trsm_gemm<F32>(s, r1 - j0 - nb, m, nb, mone, Mat{A.p + (long long)(j0 + nb) * A.ld + j0, A.ld, A.stride},
                      Mat{B.p + (long long)j0 * B.ld, B.ld, B.stride}, one,
                      Mat{B.p + (long long)(j0 + nb) * B.ld, B.ld, B.stride}, batch));
    }
    return 0;
  }
  const int mid = r0 + ((r1 - r0) / 2 + NB - 1) / NB * NB;
  KHX_TRY((trsm_lower_rec<NB, F32>(s, n, m, A, B, batch, r0, mid)));
  KHX_TRY(trsm_gemm<F32>(s, r1 - mid, m, mid - r0, mone, Mat{A.p + (long long)mid * A.ld + r0, A.ld, A.stride},
                Mat{B.p + (long long)r0 * B.ld, B.ld, B.stride}, one, Mat{B.p + (long long)mid * B.ld, B.ld, B.stride},
                batch));
  return trsm_lower_rec<NB, F32>(s, n, m, A, B, batch, mid, r1);
}

template <int NB, bool F32 = false>
static int trsm_upper_rec(hipStream_t s, int n, int m, Mat A, Mat B, int batch, int r0, int r1) {
  const cd one = mkc(1.0, 0.0), mone = mkc(-1.0, 0.0);
  if (r1 - r0 <= 2 * NB) {
    // blocks of NB rows aligned to r0, from the bottom
    const int nblk = (r1 - r0 + NB - 1) / NB;
    for (int blk = nblk - 1; blk >= 0; --blk) {
      const int j0 = r0 + blk * NB;
      const int nb = min(NB, r1 - j0);
      zlu_upper_block_kernel<NB><<<dim3((m + LB_THREADS - 1) / LB_THREADS, batch), dim3(LB_THREADS), 0, s>>>(
          min(n, j0 + nb), m, j0, A.p, A.ld, A.stride, B.p, B.ld, B.stride);
      KHX_LAUNCH_CHECK();
      if (j0 > r0)
        KHX_TRY(trsm_gemm<F32>(s, j0 - r0, m, nb, mone, Mat{A.p + (long long)r0 * A.ld + j0, A.ld, A.stride},
                      Mat{B.p + (long long)j0 * B.ld, B.ld, B.stride}, one,
                      Mat{B.p + (long long)r0 * B.ld, B.ld, B.stride}, batch));
    }
    return 0;
  }
  const int mid = r0 + ((r1 - r0) / 2 + NB - 1) / NB * NB;
  KHX_TRY((trsm_upper_rec<NB, F32>(s, n, m, A, B, batch, mid, r1)));
  KHX_TRY(trsm_gemm<F32>(s, mid - r0, m, r1 - mid, mone, Mat{A.p + (long long)r0 * A.ld + mid, A.ld, A.stride},
                Mat{B.p + (long long)mid * B.ld, B.ld, B.stride}, one, Mat{B.p + (long long)r0 * B.ld, B.ld, B.stride},
                batch));
  return trsm_upper_rec<NB, F32>(s, n, m, A, B, batch, r0, mid);
}

// Row interchanges of zgetrf (ipiv, 0-based, applied in order k = 0 .. n-1)
// to the m columns of B.  One workgroup per system: B staged in LDS when it
// fits, one thread per column walking the interchanges.
__global__ __launch_bounds__(256) void zlaswp_kernel(int n, int m, const int* __restrict__ ipiv, cd* B, int ldb,
                                                     long long sB, int staged) {
  extern __shared__ cd sb[];
  const long long item = blockIdx.x;
  cd* bm = B + item * sB;
  const int* pv = ipiv + item * n;
  if (staged) {
    for (int e = threadIdx.x; e < n * m; e += blockDim.x) sb[e] = bm[(long long)(e / m) * ldb + e % m];
    __syncthreads();
    for (int j = threadIdx.x; j < m; j += blockDim.x)
      for (int k = 0; k < n; ++k) {
        const int p = pv[k];
        if (p != k) {
          const cd t = sb[k * m + j];
          sb[k * m + j] = sb[p * m + j];
          sb[p * m + j] = t;
        }
      }
    __syncthreads();
    for (int e = threadIdx.x; e < n * m; e += blockDim.x) bm[(long long)(e / m) * ldb + e % m] = sb[e];
    return;
  }
  for (int j = threadIdx.x; j < m; j += blockDim.x)
    for (int k = 0; k < n; ++k) {
      const int p = pv[k];
      if (p != k) {
        const cd t = bm[(long long)k * ldb + j];
        bm[(long long)k * ldb + j] = bm[(long long)p * ldb + j];
        bm[(long long)p * ldb + j] = t;
      }
    }
}

int ztrsm_upper(hipStream_t s, int n, int m, Mat A, Mat B, int batch) {
  if (batch <= 0 || n <= 0 || m <= 0) return 0;
  KHX_REQUIRE(batch <= 65535, "ztrsm_upper: at most 65535 systems per call");
  return trsm_upper_rec<16>(s, n, m, A, B, batch, 0, n);
}

// B <- A^-1 B from zgetrf factors (any path): interchanges, then the two
// triangular solves split in halves into GEMMs (trsm_*_rec).
int zlu_solve_factored(hipStream_t s, int n, int m, Mat A, const int* ipiv, Mat B, int batch) {
  if (batch <= 0 || n <= 0 || m <= 0) return 0;
  KHX_REQUIRE(batch <= 65535, "zlu_solve_factored: at most 65535 systems per call");
  ProfScope ps("zgetrs_kernel", s);
  const size_t lds = (size_t)n * m * sizeof(cd);
  const int staged = lds <= 64 * 1024 ? 1 : 0;
  zlaswp_kernel<<<dim3(batch), dim3(256), staged ? lds : 0, s>>>(n, m, ipiv, B.p, B.ld, B.stride, staged);
  KHX_LAUNCH_CHECK();
  KHX_TRY(trsm_lower_rec<16>(s, n, m, A, B, batch, 0, n));
  return trsm_upper_rec<16>(s, n, m, A, B, batch, 0, n);
}

template <int NB>
static size_t lu_panel_lds(int n) { return (size_t)n * (NB + 1) * sizeof(cd); }

// One panel step (panel in LDS, or in global memory when it does not fit).
// KHX_LU_PANEL_REG=0 keeps every panel in LDS (zlu_panel_kernel)
static bool panel_reg_enabled() {
  const char* env = getenv("KHX_LU_PANEL_REG");
  return !(env && atoi(env) == 0);
}
// threads of the 16-column register panel for short panels: 64 (one
// wavefront, up to 192 rows) or 128 (up to 256 rows) through KHX_LU_PANEL_TH
static int panel_reg_threads() {
  const char* env = getenv("KHX_LU_PANEL_TH");
  return env ? atoi(env) : LB_THREADS;
}

template <int NB>
static int launch_panel(hipStream_t s, int n, int m, int j0, Mat A, Mat B, int batch, int* ipiv, int* info, int mode,
                        int cend, bool gp, size_t lds) {
  if constexpr (NB == 32) {  // one row per thread, up to 256 rows
    const int pr = n - j0;
    if (pr <= LB_THREADS && panel_reg_enabled()) {
      zlu_panel_reg_kernel<NB, 1><<<dim3(batch), dim3(LB_THREADS), NB * (NB + 1) * sizeof(cd), s>>>(
          n, m, j0, A.p, A.ld, A.stride, B.p, B.ld, B.stride, ipiv, info, mode, cend);
      KHX_LAUNCH_CHECK();
      return 0;
    }
  }
  if constexpr (NB == 16) {
    const int pr = n - j0;
    const size_t l11 = NB * (NB + 1) * sizeof(cd);
    if (pr <= 3 * 512 && panel_reg_enabled()) {
      const int th = panel_reg_threads();
      if (th == 64 && pr <= 3 * 64)
        zlu_panel_reg_kernel<NB, 3, 64><<<dim3(batch), dim3(64), l11, s>>>(n, m, j0, A.p, A.ld, A.stride, B.p, B.ld,
                                                                          B.stride, ipiv, info, mode, cend);
      else if (th == 128 && pr <= 2 * 128)
        zlu_panel_reg_kernel<NB, 2, 128><<<dim3(batch), dim3(128), l11, s>>>(n, m, j0, A.p, A.ld, A.stride, B.p,
                                                                            B.ld, B.stride, ipiv, info, mode, cend);
      else if (pr <= LB_THREADS)
        zlu_panel_reg_kernel<NB, 1><<<dim3(batch), dim3(LB_THREADS), l11, s>>>(n, m, j0, A.p, A.ld, A.stride, B.p,
                                                                              B.ld, B.stride, ipiv, info, mode, cend);
      else if (pr <= 2 * LB_THREADS)
        zlu_panel_reg_kernel<NB, 2><<<dim3(batch), dim3(LB_THREADS), l11, s>>>(n, m, j0, A.p, A.ld, A.stride, B.p,
                                                                              B.ld, B.stride, ipiv, info, mode, cend);
      else  // up to 1 536 rows: 512 threads, three rows each
        zlu_panel_reg_kernel<NB, 3, 512><<<dim3(batch), dim3(512), l11, s>>>(n, m, j0, A.p, A.ld, A.stride, B.p,
                                                                            B.ld, B.stride, ipiv, info, mode, cend);
      KHX_LAUNCH_CHECK();
      return 0;
    }
  }
  if (gp) {
    zlu_panel_kernel<NB, true><<<dim3(batch), dim3(LB_THREADS), 0, s>>>(n, m, j0, A.p, A.ld, A.stride, B.p, B.ld,
                                                                      B.stride, ipiv, info, mode, cend);
  } else {
    zlu_panel_kernel<NB, false><<<dim3(batch), dim3(LB_THREADS), lds, s>>>(n, m, j0, A.p, A.ld, A.stride, B.p, B.ld,
                                                                         B.stride, ipiv, info, mode, cend);
  }
  KHX_LAUNCH_CHECK();
  return 0;
}

// Recursive LU of the columns c0 .. c1 (rows c0 .. n), all earlier columns
// already applied:  factor the left half;  U12 = L11^-1 A12;
// A22 -= L21 U12 (one GEMM with K = half the width);  factor the right half.
// The NB-wide leaves are zlu_panel_kernel launches, which interchange every
// column (and B) at once, so no pivot list has to be kept.
template <int NB>
static int lu_rec(hipStream_t s, int n, int m, Mat A, Mat B, int batch, int* ipiv, int* info, int mode, bool gp,
                  size_t lds, int c0, int c1) {
  if (c1 - c0 <= NB) return launch_panel<NB>(s, n, m, c0, A, B, batch, ipiv, info, mode, c1, gp, lds);
  const int mid = c0 + ((c1 - c0) / 2 + NB - 1) / NB * NB;
  KHX_TRY(lu_rec<NB>(s, n, m, A, B, batch, ipiv, info, mode, gp, lds, c0, mid));
  const Mat right{A.p + mid, A.ld, A.stride};
  KHX_TRY(trsm_lower_rec<NB>(s, n, c1 - mid, A, right, batch, c0, mid));
  KHX_TRY(zgemm(s, n - mid, c1 - mid, mid - c0, mkc(-1.0, 0.0), Mat{A.p + (long long)mid * A.ld + c0, A.ld, A.stride},
                Mat{A.p + (long long)c0 * A.ld + mid, A.ld, A.stride}, mkc(1.0, 0.0),
                Mat{A.p + (long long)mid * A.ld + mid, A.ld, A.stride}, batch));
  return lu_rec<NB>(s, n, m, A, B, batch, ipiv, info, mode, gp, lds, mid, c1);
}

bool lu_multi_fits(int n) { return n >= 1; }  // larger systems keep the panel in global memory

template <int NB>
static int zlu_multi_nb(hipStream_t s, int n, int m, Mat A, Mat B, int batch, int* ipiv, int* info, bool keepL,
                        bool f32 = false) {
  if (batch <= 0 || n <= 0) return 0;
  KHX_REQUIRE(A.stride != 0 && (m == 0 || B.stride != 0) || batch == 1, "zlu_multi: operands cannot be broadcast");
  KHX_REQUIRE(batch <= 65535, "zlu_multi: at most 65535 systems per call");
  const bool gp = lu_panel_lds<NB>(n) > 150 * 1024;
  const size_t lds = gp ? 0 : lu_panel_lds<NB>(n);
  if (!gp) KHX_TRY(ensure_dynamic_lds((const void*)zlu_panel_kernel<NB, false>, lds));
  const cd one = mkc(1.0, 0.0), mone = mkc(-1.0, 0.0);
  ProfScope ps(m > 0 ? "zgesv_kernel" : "zgetrf_kernel", s);
  // right-hand sides: with 64 or more of them the panels only interchange B's
  // rows and the triangular solves run afterwards, split in halves
  // (trsm_*_rec), which needs L in final row order (interchanges applied to
  // the factored columns too); fewer right-hand sides (or KHX_LU_RECSOLVE=0)
  // are solved panel by panel
  const char* rs_env = getenv("KHX_LU_RECSOLVE");
  const bool rec = m >= 64 && !(rs_env && atoi(rs_env) == 0);
  const int mode = (keepL || rec ? 1 : 0) | (rec ? 2 : 0);
  // the factorization itself recursively (deep-K GEMM updates) for n >= 400
  // when B is absent or solved afterwards (KHX_LU_RECFACTOR=0/1 forces it)
  const char* rf_env = getenv("KHX_LU_RECFACTOR");
  // measured: with the register panels, panel-by-panel is faster below (n = 242, 64 systems: +6 %)
  const bool recf = rf_env ? atoi(rf_env) != 0 : n >= 400;
  if ((m <= 0 || rec) && recf) {
    // the TRSM / GEMM updates read L in final row order: interchanges always reach the factored columns
    KHX_TRY(lu_rec<NB>(s, n, m, A, B, batch, ipiv, info, mode | 1, gp, lds, 0, n));
    if (m <= 0) return 0;
    if (f32) {
      KHX_TRY((trsm_lower_rec<NB, true>(s, n, m, A, B, batch, 0, n)));
      return trsm_upper_rec<NB, true>(s, n, m, A, B, batch, 0, n);
    }
    KHX_TRY(trsm_lower_rec<NB>(s, n, m, A, B, batch, 0, n));
    return trsm_upper_rec<NB>(s, n, m, A, B, batch, 0, n);
  }
  for (int j0 = 0; j0 < n; j0 += NB) {
    const int nb = min(NB, n - j0);
    KHX_TRY(launch_panel<NB>(s, n, m, j0, A, B, batch, ipiv, info, mode, n, gp, lds));
    const int r1 = j0 + nb, rest = n - r1;
    if (rest > 0) {
      const Mat L21{A.p + (long long)r1 * A.ld + j0, A.ld, A.stride};
      const Mat U12{A.p + (long long)j0 * A.ld + r1, A.ld, A.stride};
      const Mat A22{A.p + (long long)r1 * A.ld + r1, A.ld, A.stride};
      KHX_TRY(zgemm(s, rest, rest, nb, mone, L21, U12, one, A22, batch));
      if (m > 0 && !rec) {
        const Mat Y1{B.p + (long long)j0 * B.ld, B.ld, B.stride};
        const Mat B2{B.p + (long long)r1 * B.ld, B.ld, B.stride};
        KHX_TRY(zgemm(s, rest, m, nb, mone, L21, Y1, one, B2, batch));
      }
    }
  }
  if (m <= 0) return 0;
  if (rec && f32) {
    KHX_TRY((trsm_lower_rec<NB, true>(s, n, m, A, B, batch, 0, n)));
    return trsm_upper_rec<NB, true>(s, n, m, A, B, batch, 0, n);
  }
  if (rec) {
    KHX_TRY(trsm_lower_rec<NB>(s, n, m, A, B, batch, 0, n));
    return trsm_upper_rec<NB>(s, n, m, A, B, batch, 0, n);
  }
  // back substitution U X = Y, NB rows at a time from the bottom
  const int nblk = (n + NB - 1) / NB;
  for (int blk = nblk - 1; blk >= 0; --blk) {
    const int j0 = blk * NB;
    const int nb = min(NB, n - j0);
    zlu_upper_block_kernel<NB><<<dim3((m + LB_THREADS - 1) / LB_THREADS, batch), dim3(LB_THREADS), 0, s>>>(
        n, m, j0, A.p, A.ld, A.stride, B.p, B.ld, B.stride);
    KHX_LAUNCH_CHECK();
    if (j0 > 0) {
      const Mat U01{A.p + j0, A.ld, A.stride};
      const Mat X1{B.p + (long long)j0 * B.ld, B.ld, B.stride};
      KHX_TRY(zgemm(s, j0, m, nb, mone, U01, X1, one, B, batch));
    }
  }
  return 0;
}

// panel width: KHX_LU_NB=16|32 forces it
int zlu_multi(hipStream_t s, int n, int m, Mat A, Mat B, int batch, int* ipiv, int* info, bool keepL) {
  const char* env = getenv("KHX_LU_NB");
  const int nb = env ? atoi(env) : 16;
  if (nb == 32) return zlu_multi_nb<32>(s, n, m, A, B, batch, ipiv, info, keepL);
  return zlu_multi_nb<16>(s, n, m, A, B, batch, ipiv, info, keepL);
}

// Mixed-precision solve for 64 or more right-hand sides (KHX_LU_MIXED=1, off
// by default: slower here, where the FP64 matrix cores run at half the FP32
// rate): the FP64 LU with partial pivoting of zlu_multi, triangular solves
// whose GEMM updates run on the FP32 matrix cores (zgemm_f32), then one FP64
// residual refinement step:
//   X0 = U^-1 L^-1 P B (FP32 products),  R = B - A X0 (FP64),  X = X0 + U^-1 L^-1 P R.
// The copies of A and B and the pivots live in a grow-only buffer of the
// library, one per device, allocated on first use (not for graph capture).
namespace {
struct RefineBuf {
  void* p = nullptr;
  size_t cap = 0;
};
std::mutex g_refine_mu;
RefineBuf g_refine[16];
}  // namespace

static int refine_buffer(size_t bytes, void** out) {
  int dev = 0;
  KHX_CHECK_HIP(hipGetDevice(&dev));
  KHX_REQUIRE(dev >= 0 && dev < 16, "zgesv_refined: device index out of range");
  std::lock_guard<std::mutex> lk(g_refine_mu);
  RefineBuf& b = g_refine[dev];
  if (b.cap < bytes) {
    if (b.p) KHX_CHECK_HIP(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    KHX_CHECK_HIP(hipMalloc(&b.p, bytes));
    b.cap = bytes;
  }
  *out = b.p;
  return 0;
}

int zgesv_refined(hipStream_t s, int n, int m, Mat A, Mat B, int batch, int* info) {
  if (batch <= 0 || n <= 0 || m <= 0) return 0;
  KHX_REQUIRE(A.ld == n && A.stride != 0 && B.stride != 0, "zgesv_refined: dense, non-broadcast systems only");
  KHX_REQUIRE(m >= 64, "zgesv_refined: needs 64 or more right-hand sides");
  const long long nn = (long long)n * n, nm = (long long)n * m;
  const size_t bytes = (size_t)batch * (nn + nm) * sizeof(cd) + (size_t)batch * n * sizeof(int) + 256;
  void* buf = nullptr;
  KHX_TRY(refine_buffer(bytes, &buf));
  cd* A0 = static_cast<cd*>(buf);
  cd* R = A0 + (long long)batch * nn;
  int* ipiv = reinterpret_cast<int*>(R + (long long)batch * nm);
  const Mat mA0{A0, n, nn}, mR{R, m, nm};
  const cd one = mkc(1.0, 0.0), mone = mkc(-1.0, 0.0);
  KHX_TRY(zcopy(s, mA0, A, n, n, batch));
  KHX_TRY(zcopy(s, mR, B, n, m, batch));
  KHX_TRY(zlu_multi_nb<16>(s, n, m, A, B, batch, ipiv, info, true, true));  // X0 in B
  KHX_TRY(zgemm(s, n, m, n, mone, mA0, B, one, mR, batch));                  // R = B - A X0
  const size_t lds = (size_t)n * m * sizeof(cd);
  const int staged = lds <= 64 * 1024 ? 1 : 0;
  zlaswp_kernel<<<dim3(batch), dim3(256), staged ? lds : 0, s>>>(n, m, ipiv, R, m, nm, staged);
  KHX_LAUNCH_CHECK();
  KHX_TRY((trsm_lower_rec<16, true>(s, n, m, A, mR, batch, 0, n)));
  KHX_TRY((trsm_upper_rec<16, true>(s, n, m, A, mR, batch, 0, n)));
  return zaxpby(s, B, one, B, one, mR, n, m, batch);
}

// KHX_LU_PATH=0/1/2 forces a path; by default the multi-launch variant takes
// the larger systems (its trailing updates fill the GPU even for few systems)
int lu_path(int n) {
  const char* env = getenv("KHX_LU_PATH");
  if (env) {
    const int p = atoi(env);
    if (p == 2 && lu_multi_fits(n)) return 2;
    if (p >= 1 && n >= 24 && lu_blocked_fits(n)) return 1;
    return 0;
  }
  if (getenv("KHX_LU_UNBLOCKED")) return 0;
  if (n >= 96 && lu_multi_fits(n)) return 2;
  if (n >= 24 && lu_blocked_fits(n)) return 1;
  return 0;
}

static size_t lu_blocked_lds(int n) { return ((size_t)n * LB_PS + (size_t)LB_NB * LB_TW) * sizeof(cd); }

bool lu_blocked_fits(int n) { return lu_blocked_lds(n) <= 150 * 1024; }

int zlu_blocked(hipStream_t s, int n, int m, Mat A, Mat B, int batch, int* ipiv, int* info, bool keepL) {
  if (batch <= 0 || n <= 0) return 0;
  const size_t lds = lu_blocked_lds(n);
  KHX_REQUIRE(lds <= 150 * 1024, "zlu_blocked: n too large for the LDS panel");
  KHX_TRY(ensure_dynamic_lds((const void*)zlu_blocked_kernel, lds));
  ProfScope ps(m > 0 ? "zgesv_kernel" : "zgetrf_kernel", s);
  zlu_blocked_kernel<<<dim3(batch), dim3(LB_THREADS), lds, s>>>(n, m, A.p, A.ld, A.stride, B.p, B.ld, B.stride, ipiv,
                                                                info, keepL ? 1 : 0);
  KHX_LAUNCH_CHECK();
  return 0;
}

}  // namespace khx
