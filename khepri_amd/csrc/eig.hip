// Batched eigen-decomposition of general complex FP64 matrices (the P*Q
// eigenproblem of a patterned layer, khepri/alternative.py:172).
//
// One workgroup per matrix:
//   1. Householder reduction to upper Hessenberg form H = Z^H A Z;
//   2. implicit single-shift complex QR (Wilkinson shift, exceptional shifts
//      every 10 sweeps, Ahues-Tisseur deflation test) to the Schur form T;
//   3. eigenvectors of the triangular T by back substitution, tiny pivots
//      replaced by max(ulp*|lambda|_1, smallnum) as LAPACK ztrevc does;
//   4. (separate launches) W = Z X, columns normalised to unit 2-norm with
//      the largest component real -- the zgeev convention.
//
// Memory layout: Z is kept transposed (row k of Zt = Schur vector k) so that
// both the reflector and the rotation updates of Z are contiguous rows.  H
// lives in LDS when it fits (n <= KHX_EIG_LDS_N), in global memory otherwise.
// In the QR phase a rotation updates only the active window [l, ihi]
// immediately (one barrier per rotation; the inputs of the next rotation
// travel through LDS); the sweep's rotations are recorded in LDS and applied
// once per sweep to the rows above the window, the columns right of it and
// to Z, each thread sliding along its own row/column.
#include "khx_internal.h"
#include "khx_prof.h"

#include <cstdlib>
#include <mutex>
#include <vector>

namespace khx {

constexpr int EIG_THREADS_DEFAULT = 256;
constexpr double ULP = 2.220446049250313e-16;
constexpr double SAFMIN = 2.2250738585072014e-308;
constexpr int KHX_EIG_LDS_N = 64;  // H resident in LDS up to this order
#ifndef EIG_QB_SIZE
#define EIG_QB_SIZE 40  // 40 x 41 padded block: the largest that keeps 4 workgroups per CU at n = 162
#endif
constexpr int EIG_QB = EIG_QB_SIZE;  // QR chase block held in LDS when H is in global memory

#ifndef HB_NB_SIZE
#define HB_NB_SIZE 8
#endif
constexpr int HB_NB = HB_NB_SIZE;  // panel width of the blocked Hessenberg reduction
#ifndef EIG_MV_ROWS
#define EIG_MV_ROWS 2  // rows per wavefront step in the Hessenberg panel's matrix-vector product (1, 3, 4, 8: slower)
#endif

// multishift QR (chain passes, see ms_chain_pass)
#ifndef MS_NS_SIZE
#define MS_NS_SIZE 8
#endif
#ifndef MS_W_SIZE
#define MS_W_SIZE 40
#endif
constexpr int MS_NS = MS_NS_SIZE;        // bulges (shifts) per chain pass
constexpr int MS_W = MS_W_SIZE;          // rows / columns of the LDS window
constexpr int MS_LAG = 2 * (MS_NS - 1);  // the trailing bulge's distance behind the leading one
constexpr int MS_R = 2 * MS_NS;          // live elements of a line around the chain
constexpr int MS_PF = 4;                 // loads in flight per line
constexpr int MS_CH = 32;                // steps of rotation parameters staged per chunk
constexpr int MS_SLOTS = MS_W * MS_NS;   // rotation slots of one window (at most MS_W steps)
constexpr int MS_MIN_WL = 3 * MS_NS;     // smaller active blocks take single-shift sweeps
static_assert(MS_W >= 2 * MS_NS + 3 && MS_W <= 64, "the chain must fit the window, the window a wavefront");
static_assert(MS_CH * MS_NS <= MS_SLOTS, "staging chunk");

// LDS layout (complex elements): region 1 = u, vh, srot (n each) and crot
// (n doubles), or the rotation slots of a chain window; wv (2n, unblocked
// Hessenberg only); part (the Hessenberg partial sums / T factor, the QR
// chase block or chain window, staged rotation parameters); then H itself
// when n <= KHX_EIG_LDS_N, else the chain pass's row table and shifts.
__host__ __device__ __forceinline__ int eig_r1_elems(int n, bool chain) {
  const int base = 3 * n + (n + 1) / 2;
  const int slots = chain && n > KHX_EIG_LDS_N ? MS_SLOTS + MS_SLOTS / 2 : 0;
  return base > slots ? base : slots;
}
__host__ __device__ __forceinline__ int eig_wv_elems(int n) { return n <= KHX_EIG_LDS_N ? 2 * n : 0; }
__host__ __device__ __forceinline__ int eig_part_elems(int n, int waves, bool chain) {
  if (n <= KHX_EIG_LDS_N) return waves * 2 * n;
  int e = HB_NB * HB_NB + waves * HB_NB + 2 * HB_NB + 2 * 8 * HB_NB;  // T, sums, staged rows
  if (e < EIG_QB * (EIG_QB + 1)) e = EIG_QB * (EIG_QB + 1);            // single-shift chase block
  if (chain) {
    if (e < MS_W * (MS_W + 1)) e = MS_W * (MS_W + 1);                  // chain window
    if (e < MS_NS * (MS_NS + 1)) e = MS_NS * (MS_NS + 1);              // shift computation
    if (e < MS_CH * MS_NS + (MS_CH * MS_NS + 1) / 2) e = MS_CH * MS_NS + (MS_CH * MS_NS + 1) / 2;  // staging
  }
  return e;
}
// The Hessenberg stage's LDS copy of the current panel (rows x HB_NB, flags
// bit 3) sits in part after the stage's own factors and staged rows.
__host__ __device__ __forceinline__ int hess_pan_offset(int waves) {
  return HB_NB * HB_NB + waves * HB_NB + 2 * HB_NB + 2 * 8 * HB_NB;
}
__host__ __device__ __forceinline__ int hess_part_elems(int n, int waves) { return hess_pan_offset(waves) + n * HB_NB; }
__host__ __device__ __forceinline__ int eig_tail_elems(int n, bool chain) {
  return n <= KHX_EIG_LDS_N ? n * n : (chain ? (n + 3) / 4 + MS_NS : 0);
}

template <int EIG_THREADS>
__device__ __forceinline__ double block_sum(double v, double* red) {
  constexpr int EIG_WAVES = EIG_THREADS / 64;
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

template <int EIG_THREADS>
__device__ __forceinline__ int block_max_int(int v, int* red) {
  constexpr int EIG_WAVES = EIG_THREADS / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  if (lane == 0) red[wid] = v;
  __syncthreads();
  int t = red[0];
#pragma unroll
  for (int w = 1; w < EIG_WAVES; ++w) t = max(t, red[w]);
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(t);  // uniform: a scalar register
}

// 1/sqrt(x) for x in the normal range: the hardware estimate refined by the
// second-order step y + y e (1/2 + 3/8 e), e = 1 - x y^2 (the series of
// (1 - e)^(-1/2)); no zero / infinity / NaN handling, which the caller's range
// test makes unnecessary.
__device__ __forceinline__ double rsqrt_normal(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Complex Givens rotation G = [[c, s], [-conj(s), c]] with G [x; y] = [r; 0].
// Fast path: two independent reciprocal square roots instead of the chain of
// hypot calls and divisions; squares stay far from over/underflow there.
__device__ __forceinline__ void givens(cd x, cd y, double& c, cd& s, cd& r) {
  const double ax2 = abs2c(x), ay2 = abs2c(y);
  const double nr2 = ax2 + ay2;
  if (ax2 > 1e-280 && nr2 < 1e280 && ay2 > 0.0) {
    const double iax = rsqrt_normal(ax2), inr = rsqrt_normal(nr2);
    const double ax = ax2 * iax, nr = nr2 * inr;
    c = ax * inr;
    s = (iax * inr) * (x * conjc(y));
    r = (nr * iax) * x;
    return;
  }
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
  const cd hkk1 = H[k * n + k - 1];
  const double h = cabs1(hkk1);
  if (h <= smlnum) return true;
  const cd a = H[(k - 1) * n + k - 1];
  const cd d = H[k * n + k];
  const double tst = cabs1(a) + cabs1(d);
  if (tst == 0.0) return false;
  if (h > ULP * tst) return false;
  const double up = cabs1(H[(k - 1) * n + k]);
  const double ab = fmax(h, up), ba = fmin(h, up);
  const double dd = cabs1(a - d), ad = cabs1(d);
  const double aa = fmax(ad, dd), bb = fmin(ad, dd);
  const double s = aa + ab;
  return ba * (ab / s) <= fmax(smlnum, ULP * (bb * (aa / s)));
}

// column rotation (H G^H on columns k, k+1): (x, y) <- (c x + conj(s) y, c y - s x)
__device__ __forceinline__ void rot_cols(cd& x, cd& y, double c, cd s) {
  const cd nx = fmac(c * x, conjc(s), y);
  y = fmsc(c * y, s, x);
  x = nx;
}
// row rotation (G H on rows k, k+1): (x, y) <- (c x + s y, c y - conj(s) x)
__device__ __forceinline__ void rot_rows(cd& x, cd& y, double c, cd s) {
  const cd nx = fmac(c * x, s, y);
  y = fmsc(c * y, conjc(s), x);
  x = nx;
}

// Apply the recorded rotations k = l..ihi-1 to the line v[l*step], ...,
// v[ihi*step] (a row of H above the window: step 1, column rotations; a
// column right of it or of Zt: step n).  Loads go out eight at a time so the
// sequential chain does not wait on each one.
template <bool ROWS, int PF = 8>
__device__ __forceinline__ void slide(cd* v, int step, int l, int ihi, const double* crot, const cd* srot) {
  cd xv = v[l * step];
  int k = l;
  for (; k + PF <= ihi; k += PF) {
    cd yv[PF];
#pragma unroll
    for (int m = 0; m < PF; ++m) yv[m] = v[(k + 1 + m) * step];
#pragma unroll
    for (int m = 0; m < PF; ++m) {
      if (ROWS)
        rot_rows(xv, yv[m], crot[k + m], srot[k + m]);
      else
        rot_cols(xv, yv[m], crot[k + m], srot[k + m]);
      v[(k + m) * step] = xv;
      xv = yv[m];
    }
  }
  for (; k < ihi; ++k) {
    cd yv = v[(k + 1) * step];
    if (ROWS)
      rot_rows(xv, yv, crot[k], srot[k]);
    else
      rot_cols(xv, yv, crot[k], srot[k]);
    v[k * step] = xv;
    xv = yv;
  }
  v[ihi * step] = xv;
}

// Blocked Householder reduction to Hessenberg form for H in global memory
// (the panel scheme of LAPACK zgehrd / zlahr2, with reflectors P = I - u u^H,
// |u|^2 = 2).  For a panel of HB_NB columns starting at k0 the reflectors
// form Q = P_0 ... P_{nb-1} = I - V T V^H, and Y = A V T is built column by
// column from one matrix-vector product per reflector.  Each panel column is
// brought up to date (right update by Y, left update by V T^H V^H) just
// before its reflector is generated; after the panel the trailing columns get
// A <- (I - V T^H V^H)(A - Y V^H) and the Schur-vector matrix Zt <- Q^T Zt,
// each as one pass per column instead of one pass per reflector.
// V and Y (n x HB_NB, row-major) live in global scratch; T, the partial sums
// and the reflector in LDS.
#ifndef HB_RC_SIZE
#define HB_RC_SIZE 4
#endif
constexpr int HB_RC = HB_RC_SIZE;  // rows per staged chunk in the trailing passes

// One pass over the columns [c0, n) of M (rows r0..n-1), one thread per column:
//   MODE 0:  M(i, col) -= sum_q F(i, q) conj(G(col, q))            (right update, F = Y, G = V)
//   MODE 1:  w_q = sum_i conj(F(i, q)) M(i, col);  M(i, col) -= sum_q G(i, q) w_q    (F = Vt, G = V)
//   MODE 2:  w_q = sum_i F(i, q) M(i, col);        M(i, col) -= sum_q conj(G(i, q)) w_q
// F and G rows (n x HB_NB, row-major) are staged HB_RC rows at a time in LDS.
template <int EIG_THREADS>
__device__ __forceinline__ void stage_rows(cd* sf, const cd* src, int i0, int n) {
  __syncthreads();  // previous chunk consumed
  for (int e = threadIdx.x; e < HB_RC * HB_NB; e += EIG_THREADS) {
    const int i = i0 + e / HB_NB;
    sf[e] = (i < n) ? src[(long long)i * HB_NB + e % HB_NB] : mkc(0.0, 0.0);
  }
  __syncthreads();
}

// M(i, col) -= sum_q S(i, q) w_q for rows [r0, n) (CONJ: conj(S)); S staged HB_RC rows at a time
template <int EIG_THREADS, bool CONJ>
__device__ __forceinline__ void update_rows(cd* M, int n, int cc, bool has, int r0, const cd* S, const cd* w,
                                            cd* sf) {
#pragma unroll 1
  for (int i0 = r0; i0 < n; i0 += HB_RC) {
    stage_rows<EIG_THREADS>(sf, S, i0, n);
    if (has) {
      cd a[HB_RC];
#pragma unroll
      for (int r = 0; r < HB_RC; ++r) a[r] = (i0 + r < n) ? M[(long long)(i0 + r) * n + cc] : mkc(0.0, 0.0);
#pragma unroll
      for (int r = 0; r < HB_RC; ++r) {
#pragma unroll
        for (int q = 0; q < HB_NB; ++q) {
          const cd g = sf[r * HB_NB + q];
          a[r] = fmsc(a[r], CONJ ? conjc(g) : g, w[q]);
        }
        if (i0 + r < n) M[(long long)(i0 + r) * n + cc] = a[r];
        __builtin_amdgcn_sched_barrier(0);  // keep the next row's LDS reads from being hoisted
      }
    }
  }
}

// w_q += sum_i S(i, q) M(i, col) over rows [r0, n) (CONJ: conj(S))
template <int EIG_THREADS, bool CONJ>
__device__ __forceinline__ void accumulate_rows(const cd* M, int n, int cc, bool has, int r0, const cd* S, cd* w,
                                                cd* sf) {
#pragma unroll 1
  for (int i0 = r0; i0 < n; i0 += HB_RC) {
    stage_rows<EIG_THREADS>(sf, S, i0, n);
    if (has) {
      cd a[HB_RC];
#pragma unroll
      for (int r = 0; r < HB_RC; ++r) a[r] = (i0 + r < n) ? M[(long long)(i0 + r) * n + cc] : mkc(0.0, 0.0);
#pragma unroll
      for (int r = 0; r < HB_RC; ++r) {
#pragma unroll
        for (int q = 0; q < HB_NB; ++q) {
          if (CONJ)
            w[q] = fmac_cj(w[q], sf[r * HB_NB + q], a[r]);
          else
            w[q] = fmac(w[q], sf[r * HB_NB + q], a[r]);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
}

// One pass over the columns [c0, n) of M, one thread per column:
//   MODE 0:  M(i, col) -= sum_q F(i, q) conj(G(col, q)), all rows         (F = Y, G = V)
//   MODE 1:  w = F^H M(r0:, col);  M(r0:, col) -= G w                      (F = Vt, G = V)
//   MODE 2:  w = F^T M(r0:, col);  M(r0:, col) -= conj(G) w
template <int EIG_THREADS, int MODE>
__device__ __forceinline__ void trailing_pass(cd* M, int n, int c0, int r0, const cd* F, const cd* G, cd* stage) {
#pragma unroll 1
  for (int cb = c0; cb < n; cb += EIG_THREADS) {
    const int cc = cb + threadIdx.x;
    const bool has = cc < n;
    cd w[HB_NB];
#pragma unroll
    for (int q = 0; q < HB_NB; ++q) w[q] = (MODE == 0 && has) ? conjc(G[(long long)cc * HB_NB + q]) : mkc(0.0, 0.0);
    if (MODE == 0) {
      update_rows<EIG_THREADS, false>(M, n, cc, has, r0, F, w, stage);
    } else {
      accumulate_rows<EIG_THREADS, MODE == 1>(M, n, cc, has, r0, F, w, stage);
      update_rows<EIG_THREADS, MODE == 2>(M, n, cc, has, r0, G, w, stage);
    }
  }
  __syncthreads();
}

// LDSP: the panel's columns (rows k0+1..) live in LDS while the panel is
// factored (steps 1-3 read and write them there; the matrix-vector products
// read the panel-start matrix from H, whose panel columns are untouched
// until the panel is written back).  Same operations in the same order.
template <int EIG_THREADS, bool LDSP>
__device__ __forceinline__ void hessenberg_blocked(cd* H, cd* Zt, int n, cd* V, cd* Y, cd* Vt, cd* u, cd* vh, cd* part,
                                   double* red) {
  constexpr int EIG_WAVES = EIG_THREADS / 64;
  static_assert(64 % HB_NB == 0 && EIG_THREADS % HB_NB == 0, "panel width must divide the wavefront");
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  cd* Ts = part;                      // HB_NB x HB_NB, zero outside the panel's reflectors
  cd* wsum = part + HB_NB * HB_NB;    // EIG_WAVES x HB_NB partial sums
  cd* wv = wsum + EIG_WAVES * HB_NB;  // HB_NB reduced vector
  cd* wt = wv + HB_NB;                // HB_NB transformed vector
  cd* stage = wt + HB_NB;             // 2 x HB_RC x HB_NB staged factor rows
  cd* pan = part + hess_pan_offset(EIG_WAVES);  // LDSP: pan[i * HB_NB + j] = H[i][k0 + j], rows k0+1..
  // element (i, c) of the panel column c = k0 + j
  auto pc = [&](int i, int c, int j) -> cd& { return LDSP ? pan[i * HB_NB + j] : H[(long long)i * n + c]; };
  // wv[q] = sum_{i >= r0} conj(V(i, q)) x_i for every q (columns of V past the
  // current reflector are zero): thread -> (q = tid % HB_NB, every 16th row).
  auto vh_dot = [&](const cd* x, long long xstride, int r0) {
    const int q = tid % HB_NB;
    cd acc = mkc(0.0, 0.0);
    for (int i = r0 + tid / HB_NB; i < n; i += EIG_THREADS / HB_NB)
      acc = fmac_cj(acc, V[(long long)i * HB_NB + q], x[(long long)i * xstride]);
#pragma unroll
    for (int o = HB_NB; o < 64; o <<= 1) {
      acc.x += __shfl_xor(acc.x, o, 64);
      acc.y += __shfl_xor(acc.y, o, 64);
    }
    if (lane < HB_NB) wsum[wid * HB_NB + lane] = acc;
    __syncthreads();
    if (tid < HB_NB) {
      cd t = wsum[tid];
#pragma unroll
      for (int w = 1; w < EIG_WAVES; ++w) t = t + wsum[w * HB_NB + tid];
      wv[tid] = t;
    }
    __syncthreads();
  };

  for (int k0 = 0; k0 + 2 < n; k0 += HB_NB) {
    const int nbp = min(HB_NB, n - 2 - k0);  // reflectors of this panel
    for (int e = tid; e < HB_NB * HB_NB; e += EIG_THREADS) Ts[e] = mkc(0.0, 0.0);
    for (int e = tid; e < n * HB_NB; e += EIG_THREADS) {
      V[e] = mkc(0.0, 0.0);
      Y[e] = mkc(0.0, 0.0);
    }
    if (LDSP)
      for (int e = (k0 + 1) * HB_NB + tid; e < n * HB_NB; e += EIG_THREADS)
        if (e % HB_NB < nbp) pan[e] = H[(long long)(e / HB_NB) * n + k0 + e % HB_NB];
    __syncthreads();
    for (int j = 0; j < nbp; ++j) {
      const int c = k0 + j;
      if (j > 0) {
        // (1) right update of column c by the panel's previous reflectors: a -= Y V(c, :)^H
        // (rows k0+1..; the rows above the panel get it from Y's top rows after the panel)
        for (int i = k0 + 1 + tid; i < n; i += EIG_THREADS) {
          const cd* yr = Y + (long long)i * HB_NB;
          const cd* vc = V + (long long)c * HB_NB;
          cd a = pc(i, c, j);
          for (int q = 0; q < j; ++q) a = fmsc(a, yr[q], conjc(vc[q]));
          pc(i, c, j) = a;
        }
        __syncthreads();
        // (2) left update of rows k0+1.. of column c: x -= V T^H V^H x
        if (LDSP)
          vh_dot(pan + j, HB_NB, k0 + 1);
        else
          vh_dot(H + c, n, k0 + 1);
        if (tid < j) {
          cd t = mkc(0.0, 0.0);
          for (int p = 0; p <= tid; ++p) t = fmac_cj(t, Ts[p * HB_NB + tid], wv[p]);
          wt[tid] = t;
        }
        __syncthreads();
        for (int i = k0 + 1 + tid; i < n; i += EIG_THREADS) {
          const cd* vr = V + (long long)i * HB_NB;
          cd a = pc(i, c, j);
          for (int q = 0; q < j; ++q) a = fmsc(a, vr[q], wt[q]);
          pc(i, c, j) = a;
        }
        __syncthreads();
      }
      // (3) reflector from H[c+1.., c]
      const int L = n - c - 1;
      const cd alpha = pc(c + 1, c, j);  // read before the barrier in block_sum
      double part2 = 0.0;
      for (int i = c + 2 + tid; i < n; i += EIG_THREADS) part2 += abs2c(pc(i, c, j));
      const double rest2 = block_sum<EIG_THREADS>(part2, red);
      const double aa = absc(alpha);
      const double xn = sqrt(aa * aa + rest2);
      const cd ph = aa != 0.0 ? (1.0 / aa) * alpha : mkc(1.0, 0.0);
      const bool skip = (rest2 == 0.0);  // already reduced: P = I (u = 0)
      const double sc = skip ? 0.0 : 1.0 / sqrt(xn * (xn + aa));
      for (int i = c + 1 + tid; i < n; i += EIG_THREADS) {
        cd x = pc(i, c, j);
        if (i == c + 1) x = x + xn * ph;
        const cd v = sc * x;
        u[i - c - 1] = v;
        V[(long long)i * HB_NB + j] = v;
        if (!skip) pc(i, c, j) = (i == c + 1) ? -(xn * ph) : mkc(0.0, 0.0);
      }
      __syncthreads();
      // (4) vh = H(:, c+1:) u  (panel-start columns): each wavefront takes
      // EIG_MV_ROWS rows at a time (as many independent loads in flight per lane)
      for (int i0 = k0 + 1 + wid * EIG_MV_ROWS; i0 < n; i0 += EIG_WAVES * EIG_MV_ROWS) {
        double sx[EIG_MV_ROWS], sy[EIG_MV_ROWS];
#pragma unroll
        for (int r = 0; r < EIG_MV_ROWS; ++r) sx[r] = sy[r] = 0.0;
        for (int q = lane; q < L; q += 64) {
          const cd uq = u[q];
#pragma unroll
          for (int r = 0; r < EIG_MV_ROWS; ++r)
            if (i0 + r < n) {
              const cd t = H[(long long)(i0 + r) * n + c + 1 + q] * uq;
              sx[r] += t.x;
              sy[r] += t.y;
            }
        }
#pragma unroll
        for (int r = 0; r < EIG_MV_ROWS; ++r) {
          const double a = wave_sum(sx[r]), b = wave_sum(sy[r]);
          if (lane == 0 && i0 + r < n) vh[i0 + r] = mkc(a, b);
        }
      }
      __syncthreads();
      // (5) z = V(:, 0:j)^H u; T(0:j, j) = -T z, T(j, j) = 1; Y(:, j) = vh - Y(:, 0:j) z
      if (j > 0) vh_dot(V + j, HB_NB, c + 1);
      if (tid < j) {
        cd t = mkc(0.0, 0.0);
        for (int p = tid; p < j; ++p) t = fmsc(t, Ts[tid * HB_NB + p], wv[p]);
        Ts[tid * HB_NB + j] = t;
      } else if (tid == j) {
        Ts[j * HB_NB + j] = mkc(skip ? 0.0 : 1.0, 0.0);
      }
      for (int i = k0 + 1 + tid; i < n; i += EIG_THREADS) {
        const cd* yr = Y + (long long)i * HB_NB;
        cd a = vh[i];
        for (int q = 0; q < j; ++q) a = fmsc(a, yr[q], wv[q]);
        Y[(long long)i * HB_NB + j] = a;
      }
      __syncthreads();
    }
    if (LDSP)  // the finished panel columns back to H (their rows 0..k0 are updated in H below)
      for (int e = (k0 + 1) * HB_NB + tid; e < n * HB_NB; e += EIG_THREADS)
        if (e % HB_NB < nbp) H[(long long)(e / HB_NB) * n + k0 + e % HB_NB] = pan[e];
    // Vt = V T (the left factor of both trailing updates), one thread per row
    for (int i = tid; i < n; i += EIG_THREADS) {
      const cd* vr = V + (long long)i * HB_NB;
      cd* tr = Vt + (long long)i * HB_NB;
#pragma unroll 1
      for (int q = 0; q < HB_NB; ++q) {
        cd t = mkc(0.0, 0.0);
#pragma unroll 1
        for (int p = 0; p <= q; ++p) t = fmac(t, vr[p], Ts[p * HB_NB + q]);
        tr[q] = t;
      }
    }
    __syncthreads();
    // Y's rows 0..k0 in one pass (zlahr2's Y(1:K) = A(1:K, K+1:N) V T): the
    // columns past k0 are still at their panel-start values in those rows
    for (int e = tid; e < (k0 + 1) * HB_NB; e += EIG_THREADS) {
      const int i = e / HB_NB, q = e % HB_NB;
      const cd* hr = H + (long long)i * n;
      cd acc = mkc(0.0, 0.0);
      for (int m = k0 + 1; m < n; ++m) acc = fmac(acc, hr[m], Vt[(long long)m * HB_NB + q]);
      Y[(long long)i * HB_NB + q] = acc;
    }
    __syncthreads();
    // rows 0..k0 of the panel's later columns: a -= Y V(c, :)^H
    for (int e = tid; e < (k0 + 1) * (nbp - 1); e += EIG_THREADS) {
      const int i = e / (nbp - 1), c = k0 + 1 + e % (nbp - 1);
      const cd* yr = Y + (long long)i * HB_NB;
      const cd* vc = V + (long long)c * HB_NB;
      cd a = H[(long long)i * n + c];
      for (int q = 0; q < HB_NB; ++q) a = fmsc(a, yr[q], conjc(vc[q]));
      H[(long long)i * n + c] = a;
    }
    __syncthreads();
    // (6)-(8) trailing columns A <- (I - V Vt^H)(A - Y V^H) and Zt <- Zt - conj(V) Vt^T Zt,
    // one thread per column; the factor rows (Y, V, Vt) of HB_RC rows at a time
    // are staged in LDS and the column's HB_RC elements are loaded together.
    trailing_pass<EIG_THREADS, 0>(H, n, k0 + nbp, 0, Y, V, stage);       // A(:, trail) -= Y V(trail, :)^H
    trailing_pass<EIG_THREADS, 1>(H, n, k0 + nbp, k0 + 1, Vt, V, stage);  // A -= V (Vt^H A)  (rows k0+1..)
    trailing_pass<EIG_THREADS, 2>(Zt, n, 0, k0 + 1, Vt, V, stage);        // Zt -= conj(V) (Vt^T Zt)
    __syncthreads();
  }
}

// Eigenvectors of the upper-triangular Schur factor T (global memory), one
// thread per eigenvector: column k of X has X[k][k] = 1, zeros below, and
// for i < k, from the bottom up,
//   X[i][k] = (-T[i][k] - sum_{i<m<k} T[i][m] X[m][k]) / (T[i][i] - T[k][k])
// (ztrevc's back substitution, with its small-pivot perturbation).  Rows are
// taken in aligned blocks of EV_NB from the bottom: the block's rows live in
// registers while they are finished, then every row above takes the block's
// contributions in one read-modify-write.  The operations on each element
// are those of the row-by-row sweep and in the same order (subtractions in
// decreasing m, the division last), so the results are identical to it; the
// traffic is one pass over the rows above per block instead of one per row,
// and no barrier is needed (a thread only touches its own column).
constexpr int EV_NB = 16;
template <int EIG_THREADS>
__device__ __forceinline__ void eigvec_columns(const cd* __restrict__ T, cd* __restrict__ X, int n, double smlnum) {
  constexpr int EIG_WAVES = EIG_THREADS / 64;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  // rows k.. of every column: X[k][k] = 1, zeros below
  for (int j = wid; j < n; j += EIG_WAVES)
    for (int kk = lane; kk <= j; kk += 64) X[(long long)j * n + kk] = mkc(j == kk ? 1.0 : 0.0, 0.0);
  for (int k0w = wid * 64; k0w < n; k0w += EIG_THREADS) {  // a wavefront takes columns k0 .. k0 + 63
    const int k0 = __builtin_amdgcn_readfirstlane(k0w);
    const int k = k0 + lane;
    const bool valid = k < n;
    const cd lk = valid ? T[(long long)k * n + k] : mkc(0.0, 0.0);
    const double smin = fmax(ULP * cabs1(lk), smlnum);
    const int kmax = min(n - 1, k0 + 63);
    const int kfirst = valid && k >= 1 ? (k - 1) / EV_NB : -1;  // this column's first (lowest) block
    for (int b = kmax >= 1 ? (kmax - 1) / EV_NB : -1; b >= 0; --b) {  // uniform over the wavefront
      const int r0 = b * EV_NB;
      const bool act = valid && r0 < k;
      const bool first = b == kfirst;
      // the block's rows; rows k.. (in the column's first block) stay 0, so
      // that they change nothing in the updates of the rows above (x - t 0 = x)
      cd x[EV_NB];
#pragma unroll
      for (int r = 0; r < EV_NB; ++r) {
        const int j = r0 + r;
        x[r] = mkc(0.0, 0.0);
        if (act && j < k) {
          const cd t = first ? T[(long long)j * n + k] : X[(long long)j * n + k];
          x[r] = first ? -t : t;
        }
      }
#pragma unroll
      for (int r = EV_NB - 1; r >= 0; --r) {
        const int j = r0 + r;
        if (act && j < k) {
          cd d = T[(long long)j * n + j] - lk;
          if (cabs1(d) < smin) d = mkc(smin, 0.0);
          x[r] = divc(x[r], d);
#pragma unroll
          for (int i = 0; i < r; ++i) x[i] = fmsc(x[i], T[(long long)(r0 + i) * n + j], x[r]);
        }
      }
#pragma unroll
      for (int r = 0; r < EV_NB; ++r)
        if (act && r0 + r < k) X[(long long)(r0 + r) * n + k] = x[r];
      // rows above the block (a uniform loop: row i of T's block columns is the
      // same for every lane)
      for (int i = 0; i < r0; ++i) {
        const cd* ti = T + (long long)i * n + r0;
        cd v = mkc(0.0, 0.0);
        if (act) {
          const cd t = first ? T[(long long)i * n + k] : X[(long long)i * n + k];
          v = first ? -t : t;
        }
#pragma unroll
        for (int r = EV_NB - 1; r >= 0; --r) v = fmsc(v, ti[r], x[r]);
        if (act) X[(long long)i * n + k] = v;
      }
    }
  }
}

// Deferred rotations of up to EIG_KLOG QR sweeps, applied in one pass per line.
// Sweep q rotated the pairs (p, p+1), p in [lo_q, hi_q - 1]; its rotation at p
// must follow the previous sweep's at p - 1 .. p + 1, so at step t the pass
// applies sweep q at p = t - 2q: a line is read and written once for all the
// logged sweeps, its live part (2 EIG_KLOG elements) staying in registers.
#ifndef EIG_KLOG_N
#define EIG_KLOG_N 6  // sweeps per deferred pass (4: +3 % eigensolver time; 7, 8: same as 6)
#endif
constexpr int EIG_KLOG = EIG_KLOG_N;
constexpr int EIG_PF = 4;  // loads in flight per line in the pipelined pass

// Lines are processed EIG_THREADS at a time; the rotation parameters of the
// next EIG_CH steps are staged in LDS (stage, at least 1.5 EIG_KLOG (EIG_CH + 6)
// complex) so no step waits on a global load for them.
constexpr int EIG_CH = 32;
// the staged sines and cosines (K (EIG_CH + LAG) complex + as many doubles) live
// in the chase-block region of LDS
static_assert(3 * EIG_KLOG * (EIG_CH + 2 * (EIG_KLOG - 1)) <= 2 * EIG_QB * (EIG_QB + 1),
              "rotation staging does not fit the chase-block LDS region");

template <int EIG_THREADS>
__device__ __forceinline__ void flush_log(cd* H, cd* Zt, int n, int nsw, const int* sl, const int* sh, const double* lc,
                          const cd* ls, cd* stage) {
  constexpr int K = EIG_KLOG, W = 2 * EIG_KLOG, LAG = 2 * (EIG_KLOG - 1), CHP = EIG_CH + LAG;
  cd* st_s = stage;
  double* st_c = reinterpret_cast<double*>(stage + K * CHP);
  int pmin = n, pend = 0, lmax = 0, hmin = n;
  for (int q = 0; q < nsw; ++q) {
    pmin = min(pmin, sl[q]);
    pend = max(pend, sh[q]);
    lmax = max(lmax, sl[q]);
    hmin = min(hmin, sh[q]);
  }
  const int nrows = lmax, ncols = n - 1 - hmin;
  const int nlines = n + nrows + ncols;
  const int tend = pend + LAG;
  for (int base = 0; base < nlines; base += EIG_THREADS) {
    // this thread's line: a column of Zt (every sweep), a row of H above a
    // sweep's window (column rotations) or a column of H right of it (row rotations)
    const int e = base + threadIdx.x;
    cd* v = Zt;
    int step = n;
    bool rows = false;
    unsigned mask = 0;
    if (e < n) {
      v = Zt + e;
      mask = (1u << nsw) - 1u;
    } else if (e < n + nrows) {
      const int i = e - n;
      v = H + (long long)i * n;
      step = 1;
      for (int q = 0; q < nsw; ++q)
        if (i < sl[q]) mask |= 1u << q;
    } else if (e < nlines) {
      const int j = hmin + 1 + (e - n - nrows);
      v = H + j;
      rows = true;
      for (int q = 0; q < nsw; ++q)
        if (j > sh[q]) mask |= 1u << q;
    }
    int l_[K], h_[K];
    int lp = pend + 1, le = pmin - 1;  // the line's own range: elements its sweeps rotate
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const bool on = q < nsw && ((mask >> q) & 1u);
      l_[q] = on ? sl[q] : 1;
      h_[q] = on ? sh[q] : 0;
      if (on) {
        lp = min(lp, l_[q]);
        le = max(le, h_[q]);
      }
    }
    // r[i] holds element t - LAG + i at step t; shifted by one per step
    cd r[W];
#pragma unroll
    for (int i = 0; i < W; ++i) r[i] = mkc(0.0, 0.0);
    if (pmin >= lp && pmin <= le) r[W - 2] = v[(long long)pmin * step];
    cd pf[EIG_PF];
#pragma unroll
    for (int i = 0; i < EIG_PF; ++i) {
      const int el = pmin + 1 + i;
      pf[i] = (el >= lp && el <= le) ? v[(long long)el * step] : mkc(0.0, 0.0);
    }
    for (int t0 = pmin; t0 <= tend; t0 += EIG_CH) {
      __syncthreads();  // the previous chunk's parameters are no longer read
      for (int idx = threadIdx.x; idx < K * CHP; idx += EIG_THREADS) {
        const int q = idx / CHP, p = t0 - LAG + idx % CHP;
        const bool ok = q < nsw && p >= 0 && p < n;
        st_s[idx] = ok ? ls[q * n + p] : mkc(0.0, 0.0);
        st_c[idx] = ok ? lc[q * n + p] : 0.0;
      }
      __syncthreads();
      const int t1 = min(t0 + EIG_CH - 1, tend);
      for (int t = t0; t <= t1; ++t) {
        r[W - 1] = pf[0];
#pragma unroll
        for (int i = 0; i + 1 < EIG_PF; ++i) pf[i] = pf[i + 1];
        {
          const int el = t + 1 + EIG_PF;
          pf[EIG_PF - 1] = (el >= lp && el <= le) ? v[(long long)el * step] : mkc(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < K; ++q) {
          const int p = t - 2 * q;
          if (p >= l_[q] && p < h_[q]) {
            const int idx = q * CHP + (p - t0 + LAG);
            const double c = st_c[idx];
            const cd sn = st_s[idx];
            const int a = 2 * (K - 1 - q);  // constant after unrolling
            if (rows)
              rot_rows(r[a], r[a + 1], c, sn);
            else
              rot_cols(r[a], r[a + 1], c, sn);
          }
        }
        const int eo = t - LAG;
        if (eo >= lp && eo <= le) v[(long long)eo * step] = r[0];
#pragma unroll
        for (int i = 0; i + 1 < W; ++i) r[i] = r[i + 1];
      }
    }
  }
  __syncthreads();
}

// One implicit single-shift QR sweep over the window [l, ihi] with H in
// global memory.  The bulge is chased through EIG_QB x EIG_QB diagonal blocks
// staged in LDS: inside a block every rotation updates only the block (one
// barrier per rotation, LDS latency); the block's rotations are logged and,
// once the block is written back, applied to the rows above it, the columns
// right of it and to Zt -- each thread sliding along one line.  Consecutive
// blocks overlap by three indices so the next rotation finds its inputs
// (the bulge) inside the new block.  Every element receives the same
// rotations in the same order as an unblocked sweep.
template <int EIG_THREADS>
__device__ __forceinline__ void qr_sweep_blocked(cd* H, cd* Zt, int n, int l, int ihi, cd shift, cd* hb, double* crot, cd* srot,
                                 cd* sxy, bool one_wave, unsigned long long* acc) {
  const int tid = threadIdx.x;
  int kstart = l, b0 = l;
  while (true) {
    const int b1 = min(b0 + EIG_QB - 1, ihi);
    const int nb = b1 - b0 + 1;
    const int kend = (b1 == ihi) ? ihi - 1 : b1 - 2;  // last rotation whose bulge stays in the block
    for (int e = tid; e < nb * nb; e += EIG_THREADS) hb[e] = H[(long long)(b0 + e / nb) * n + b0 + e % nb];
    __syncthreads();
    if (tid == 0) {
      cd* xy = sxy + 2 * (kstart & 1);
      if (kstart == l) {  // b0 == l
        xy[0] = hb[0] - shift;
        xy[1] = hb[nb];
      } else {  // b0 == kstart - 1
        xy[0] = hb[nb];
        xy[1] = hb[2 * nb];
      }
    }
    __syncthreads();
    // The chase inside the block runs on one wavefront when one_wave is set:
    // its tasks (at most 2 EIG_QB - 3 per rotation) fit in the 64 lanes, and
    // lanes hand values to each other through LDS in program order, so no
    // workgroup barrier is needed per rotation; the other wavefronts wait.
    const int cw = one_wave ? 64 : EIG_THREADS;
    unsigned long long t_c0 = acc ? wall_clock64() : 0;
    for (int k = kstart; k <= kend && (!one_wave || tid < 64); ++k) {
      const cd x = sxy[2 * (k & 1)], y = sxy[2 * (k & 1) + 1];
      cd* nxt = sxy + 2 * ((k + 1) & 1);
      double c;
      cd s, r;
      givens(x, y, c, s, r);
      if (tid == 0) {
        crot[k] = c;
        srot[k] = s;
      }
      const int kk = k - b0;                                   // block coordinates
      const int nrow = b1 - k - 1;                             // row part: block columns kk+2 .. nb-1
      const int ncol = kk + ((k + 2 <= b1) ? 1 : 0);           // col part: block rows 0 .. kk-1 (+ kk+2)
      const int total = 1 + nrow + ncol;
      for (int e = tid; e < total; e += cw) {
        if (e == 0) {
          if (k > l) {
            hb[kk * nb + kk - 1] = r;
            hb[(kk + 1) * nb + kk - 1] = mkc(0.0, 0.0);
          }
          cd* p0 = hb + kk * nb + kk;
          cd* p1 = hb + (kk + 1) * nb + kk;
          cd b00 = p0[0], b01 = p0[1], b10 = p1[0], b11 = p1[1];
          rot_rows(b00, b10, c, s);
          rot_rows(b01, b11, c, s);
          rot_cols(b00, b01, c, s);
          rot_cols(b10, b11, c, s);
          p0[0] = b00;
          p0[1] = b01;
          p1[0] = b10;
          p1[1] = b11;
          nxt[0] = b10;
        } else if (e <= nrow) {
          const int j = kk + 1 + e;
          cd xk = hb[kk * nb + j], yk = hb[(kk + 1) * nb + j];
          rot_rows(xk, yk, c, s);
          hb[kk * nb + j] = xk;
          hb[(kk + 1) * nb + j] = yk;
        } else {
          const int q = e - nrow - 1;
          const int i = (q < kk) ? q : kk + 2;
          cd* pr = hb + i * nb + kk;
          cd xk = pr[0], yk = pr[1];
          rot_cols(xk, yk, c, s);
          pr[0] = xk;
          pr[1] = yk;
          if (i == kk + 2) nxt[1] = xk;
        }
      }
      if (one_wave) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      } else {
        __syncthreads();
      }
    }
    if (one_wave) __syncthreads();
    if (acc) acc[1] += wall_clock64() - t_c0;
    for (int e = tid; e < nb * nb; e += EIG_THREADS) H[(long long)(b0 + e / nb) * n + b0 + e % nb] = hb[e];
    // deferred inside the window: rows l..b0-1 above the block, columns
    // b1+1..ihi right of it (the rest of H and Z wait for flush_log)
    const int nabove = b0 - l, nright = ihi - b1;
    const int tot = nabove + nright;
    for (int e = tid; e < tot; e += EIG_THREADS) {
      if (e < nabove)
        slide<false>(H + (long long)(l + e) * n, 1, kstart, kend + 1, crot, srot);
      else
        slide<true>(H + b1 + 1 + (e - nabove), n, kstart, kend + 1, crot, srot);
    }
    __syncthreads();
    if (kend == ihi - 1) break;
    kstart = kend + 1;
    b0 = kstart - 1;
  }
}

// The same sweep with the bulge neighbourhood in registers.  In a block only
// a few entries feed the next rotation: the 2 x 2 core at rows/columns
// (k, k+1), the subdiagonal entry below it and entry (k, k+1), which depends
// on every row rotation applied to column k+1 so far.  One wavefront runs the
// chase; every lane computes the rotations and the core identically (uniform
// values, no divergence), and lane j owns block column j: it carries the
// entry of its column on the current row as a register and applies each row
// rotation to it as soon as it exists, writing the finished entry of the row
// above.  Entry (k+1, k+2) reaches the core by a cross-lane read from lane
// k+2.  Row and column rotations commute (G H G^H = (G H) G^H), so the
// column rotations of the rows above the core are applied afterwards, one
// lane per row sliding through the block's rotations.  The rest (rows above
// the block, columns right of it, Zt and beyond the window) is as in
// qr_sweep_blocked.  Block rows are padded by one element so the per-row
// slides hit distinct LDS banks.
__device__ __forceinline__ double readlane_d(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

#ifndef EIG_HESS_LOAD
#define EIG_HESS_LOAD 1  // the chase blocks move only the Hessenberg band (column >= row - 2) between H and LDS (0: whole blocks)
#endif
#ifndef EIG_SLIDE_PF
#define EIG_SLIDE_PF 8  // loads in flight per line in the in-window updates (16: register spills, slower)
#endif
// One block of the register chase on wavefront 0 (lane j): the rotations
// kstart..kend of a single-shift sweep inside the LDS copy hb (nb x nb, row
// stride nb + 1) of H[b0.., b0..], recorded in crot / srot; the block rows
// above the core take their column rotations at the end.
__device__ __forceinline__ void reg_chase_block(cd* hb, int nb, int b0, int l, int ihi, int kstart, int kend,
                                                cd shift, double* crot, cd* srot, int j, int ld) {
  int q = kstart - b0;  // block index of the current rotation: 0 in the first block, 1 after
  cd x, y;
  if (kstart == l) {
    x = hb[0] - shift;
    y = hb[ld];
  } else {
    x = hb[q * ld + q - 1];
    y = hb[(q + 1) * ld + q - 1];
  }
  cd a00 = hb[q * ld + q], a01 = hb[q * ld + q + 1];
  cd a10 = hb[(q + 1) * ld + q], a11 = hb[(q + 1) * ld + q + 1];
  const bool mine0 = (j >= q + 2 && j < nb);
  cd v = mine0 ? hb[q * ld + j] : mkc(0.0, 0.0);
  cd yj = mine0 ? hb[(q + 1) * ld + j] : mkc(0.0, 0.0);
  cd wq = (q + 2 < nb) ? hb[(q + 2) * ld + q + 1] : mkc(0.0, 0.0);
  cd dq = (q + 2 < nb) ? hb[(q + 2) * ld + q + 2] : mkc(0.0, 0.0);
  for (int k = kstart; k <= kend; ++k, ++q) {
    double c;
    cd s, r;
    givens(x, y, c, s, r);
    // prefetch the next step's untouched inputs
    const bool mine = (j >= q + 2 && j < nb);
    const bool more = (k < kend);
    cd yj_n = mkc(0.0, 0.0), wq_n = mkc(0.0, 0.0), dq_n = mkc(0.0, 0.0);
    if (more) {
      if (j >= q + 3 && j < nb) yj_n = hb[(q + 2) * ld + j];
      if (q + 3 < nb) {
        wq_n = hb[(q + 3) * ld + q + 2];
        dq_n = hb[(q + 3) * ld + q + 3];
      }
    }
    if (mine) {  // row rotation of the columns right of the core
      hb[q * ld + j] = fmac(c * v, s, yj);
      v = fmsc(c * yj, conjc(s), v);
    }
    rot_rows(a00, a10, c, s);
    rot_rows(a01, a11, c, s);
    rot_cols(a00, a01, c, s);
    rot_cols(a10, a11, c, s);
    // row k+2: (k+2, k) was zero, (k+2, k+1) is untouched so far
    const cd bulge = conjc(s) * wq;
    const cd w2 = c * wq;
    if (j == 0) {
      crot[k] = c;
      srot[k] = s;
      if (k > l) {
        hb[q * ld + q - 1] = r;
        hb[(q + 1) * ld + q - 1] = mkc(0.0, 0.0);
      }
      hb[q * ld + q] = a00;
      hb[q * ld + q + 1] = a01;
    }
    if (!more) {
      if (j == 0) {
        hb[(q + 1) * ld + q] = a10;
        hb[(q + 1) * ld + q + 1] = a11;
        if (q + 2 < nb) {
          hb[(q + 2) * ld + q] = bulge;
          hb[(q + 2) * ld + q + 1] = w2;
        }
      }
      if (mine) hb[(q + 1) * ld + j] = v;
      break;
    }
    x = a10;
    y = bulge;
    a00 = a11;
    a01 = mkc(readlane_d(v.x, q + 2), readlane_d(v.y, q + 2));
    a10 = w2;
    a11 = dq;
    yj = yj_n;
    wq = wq_n;
    dq = dq_n;
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  // column rotations of the block rows above the core: row i takes the
  // rotations k >= b0 + i + 1 (columns k - b0, k - b0 + 1)
  const int k0 = max(kstart, b0 + j + 1);
  if (j < nb && k0 <= kend) {
    cd* row = hb + j * ld - b0;  // row[col] = block entry (j, col - b0)
    cd xv = row[k0];
    int k = k0;
    for (; k + 4 <= kend + 1; k += 4) {
      cd yv[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) yv[m] = row[k + 1 + m];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        rot_cols(xv, yv[m], crot[k + m], srot[k + m]);
        row[k + m] = xv;
        xv = yv[m];
      }
    }
    for (; k <= kend; ++k) {
      cd yv = row[k + 1];
      rot_cols(xv, yv, crot[k], srot[k]);
      row[k] = xv;
      xv = yv;
    }
    row[kend + 1] = xv;
  }
}

// Single-shift sweep over H[l..ihi] through LDS blocks of EIG_QB rows.
// Wavefront 0 chases the bulge through block b + 1 while the other
// wavefronts apply the rotations of block b to the rest of the window (the
// rows above it and the columns right of block b + 1); the few columns that
// block b + 1 holds get block b's row rotations first.  Every element sees
// the rotations in sweep order: a block's rest update finishes (barrier)
// before the next block's rest update or load touches the same lines.
template <int EIG_THREADS>
__device__ __forceinline__ void qr_sweep_reg(cd* H, int n, int l, int ihi, cd shift, cd* hb, double* crot, cd* srot,
                             unsigned long long* acc) {
  const int tid = threadIdx.x;
  int kstart = l, b0 = l;
  int b1 = min(b0 + EIG_QB - 1, ihi);
  int kend = (b1 == ihi) ? ihi - 1 : b1 - 2;  // last rotation whose bulge stays in the block
  {
    const int nb = b1 - b0 + 1, ld = nb + 1;
    for (int e = tid; e < nb * nb; e += EIG_THREADS)
      if (EIG_HESS_LOAD == 0 || e % nb >= e / nb - 2) hb[(e / nb) * ld + e % nb] = H[(long long)(b0 + e / nb) * n + b0 + e % nb];
    __syncthreads();
    unsigned long long t_c0 = acc ? wall_clock64() : 0;
    if (tid < 64) reg_chase_block(hb, nb, b0, l, ihi, kstart, kend, shift, crot, srot, tid, ld);
    __syncthreads();
    if (acc) acc[1] += wall_clock64() - t_c0;
  }
  while (true) {
    const int nb = b1 - b0 + 1, ld = nb + 1;
    for (int e = tid; e < nb * nb; e += EIG_THREADS)
      if (EIG_HESS_LOAD == 0 || e % nb >= e / nb - 2) H[(long long)(b0 + e / nb) * n + b0 + e % nb] = hb[(e / nb) * ld + e % nb];
    const int nabove = b0 - l;
    if (kend == ihi - 1) {  // last block (b1 == ihi): only the rows above remain
      for (int e = tid; e < nabove; e += EIG_THREADS)
        slide<false, EIG_SLIDE_PF>(H + (long long)(l + e) * n, 1, kstart, kend + 1, crot, srot);
      __syncthreads();
      break;
    }
    // next block: rows / columns b0n..b1n
    const int b0n = kend, b1n = min(b0n + EIG_QB - 1, ihi);
    const int kstartn = kend + 1, kendn = (b1n == ihi) ? ihi - 1 : b1n - 2;
    // columns b1+1..b1n (held by the next block) take this block's row rotations now
    for (int e = tid; e < b1n - b1; e += EIG_THREADS)
      slide<true, EIG_SLIDE_PF>(H + b1 + 1 + e, n, kstart, kend + 1, crot, srot);
    __syncthreads();
    const int nbn = b1n - b0n + 1, ldn = nbn + 1;
    for (int e = tid; e < nbn * nbn; e += EIG_THREADS)
      if (EIG_HESS_LOAD == 0 || e % nbn >= e / nbn - 2)
        hb[(e / nbn) * ldn + e % nbn] = H[(long long)(b0n + e / nbn) * n + b0n + e % nbn];
    __syncthreads();
    unsigned long long t_c0 = acc ? wall_clock64() : 0;
    if (tid < 64) reg_chase_block(hb, nbn, b0n, l, ihi, kstartn, kendn, shift, crot, srot, tid, ldn);
    if (EIG_THREADS == 64 || tid >= 64) {
      // rest of this block: rows l..b0-1 (column rotations) and columns b1n+1..ihi (row rotations);
      // a one-wavefront workgroup does it after the chase
      constexpr int RW = EIG_THREADS == 64 ? 64 : EIG_THREADS - 64;
      const int nright = ihi - b1n, tot = nabove + nright;
      for (int e = EIG_THREADS == 64 ? tid : tid - 64; e < tot; e += RW) {
        if (e < nabove)
          slide<false, EIG_SLIDE_PF>(H + (long long)(l + e) * n, 1, kstart, kend + 1, crot, srot);
        else
          slide<true, EIG_SLIDE_PF>(H + b1n + 1 + (e - nabove), n, kstart, kend + 1, crot, srot);
      }
    }
    __syncthreads();
    if (acc) acc[1] += wall_clock64() - t_c0;
    kstart = kstartn;
    kend = kendn;
    b0 = b0n;
    b1 = b1n;
  }
}

// ------------------------------------------------------------ paired sweeps
// Two matrices per workgroup, one chase wavefront for both: every lane of the
// register chase evaluates the whole rotation and 2 x 2 core, so a wavefront
// whose lanes 0-31 chase one matrix's block and lanes 32-63 the other's
// issues one instruction stream for two chases.  Blocks are EIG_PQB wide (a
// half-wavefront's lanes); the two matrices keep their own windows, shifts,
// deflations and rotation logs.
constexpr int EIG_PQB = 31;  // chase block of the paired sweeps (four 256-thread workgroups per CU at n = 162)
__host__ __device__ __forceinline__ int eig_pair_elems(int n) {  // LDS per matrix: srot, crot, block
  return n + (n + 1) / 2 + EIG_PQB * (EIG_PQB + 1);
}
static_assert(3 * EIG_KLOG * (EIG_CH + 2 * (EIG_KLOG - 1)) <= 2 * EIG_PQB * (EIG_PQB + 1),
              "rotation staging does not fit a paired chase block");

// reg_chase_block for the half-wavefront of lane `lane` (on: this half has a
// block to chase).  The steps run on a counter t shared by the whole
// wavefront; a half with fewer rotations idles for the rest.  The entry
// (k+1, k+2) comes from the half's lane q+2 by readlane at an index uniform
// per half.  Per element, the operations and their order are those of
// reg_chase_block.
__device__ __forceinline__ void reg_chase_pair(bool on, cd* hb, int nb, int b0, int l, int kstart, int kend, cd shift,
                                               double* crot, cd* srot, int lane, int ld) {
  const int j = lane & 31, half = lane & 32;
  const int len = on ? kend - kstart + 1 : 0;
  const int q0 = on ? kstart - b0 : 0;
  const int tmax = max(__builtin_amdgcn_readlane(len, 0), __builtin_amdgcn_readlane(len, 32));
  const int qa = __builtin_amdgcn_readlane(q0, 0), qb = __builtin_amdgcn_readlane(q0, 32);
  int q = q0;
  cd x = mkc(0.0, 0.0), y = x, a00 = x, a01 = x, a10 = x, a11 = x, v = x, yj = x, wq = x, dq = x;
  if (len > 0) {
    if (kstart == l) {
      x = hb[0] - shift;
      y = hb[ld];
    } else {
      x = hb[q * ld + q - 1];
      y = hb[(q + 1) * ld + q - 1];
    }
    a00 = hb[q * ld + q];
    a01 = hb[q * ld + q + 1];
    a10 = hb[(q + 1) * ld + q];
    a11 = hb[(q + 1) * ld + q + 1];
    const bool mine0 = (j >= q + 2 && j < nb);
    if (mine0) {
      v = hb[q * ld + j];
      yj = hb[(q + 1) * ld + j];
    }
    if (q + 2 < nb) {
      wq = hb[(q + 2) * ld + q + 1];
      dq = hb[(q + 2) * ld + q + 2];
    }
  }
  for (int t = 0; t < tmax; ++t, ++q) {
    // lanes of entry (k+1, k+2) for the next step, uniform per half (clamped
    // into the half: a half past its last step never uses its value)
    const int sa = min(qa + t + 2, 31), sb = 32 + min(qb + t + 2, 31);
    if (t < len) {
      const int k = kstart + t;
      double c;
      cd s, r;
      givens(x, y, c, s, r);
      const bool mine = (j >= q + 2 && j < nb);
      const bool more = (t + 1 < len);
      cd yj_n = mkc(0.0, 0.0), wq_n = mkc(0.0, 0.0), dq_n = mkc(0.0, 0.0);
      if (more) {
        if (j >= q + 3 && j < nb) yj_n = hb[(q + 2) * ld + j];
        if (q + 3 < nb) {
          wq_n = hb[(q + 3) * ld + q + 2];
          dq_n = hb[(q + 3) * ld + q + 3];
        }
      }
      if (mine) {
        hb[q * ld + j] = fmac(c * v, s, yj);
        v = fmsc(c * yj, conjc(s), v);
      }
      rot_rows(a00, a10, c, s);
      rot_rows(a01, a11, c, s);
      rot_cols(a00, a01, c, s);
      rot_cols(a10, a11, c, s);
      const cd bulge = conjc(s) * wq;
      const cd w2 = c * wq;
      if (j == 0) {
        crot[k] = c;
        srot[k] = s;
        if (k > l) {
          hb[q * ld + q - 1] = r;
          hb[(q + 1) * ld + q - 1] = mkc(0.0, 0.0);
        }
        hb[q * ld + q] = a00;
        hb[q * ld + q + 1] = a01;
      }
      if (!more) {
        if (j == 0) {
          hb[(q + 1) * ld + q] = a10;
          hb[(q + 1) * ld + q + 1] = a11;
          if (q + 2 < nb) {
            hb[(q + 2) * ld + q] = bulge;
            hb[(q + 2) * ld + q + 1] = w2;
          }
        }
        if (mine) hb[(q + 1) * ld + j] = v;
      } else {
        x = a10;
        y = bulge;
        a00 = a11;
        const double ax = readlane_d(v.x, sa), ay = readlane_d(v.y, sa);
        const double bx = readlane_d(v.x, sb), by = readlane_d(v.y, sb);
        a01 = half ? mkc(bx, by) : mkc(ax, ay);
        a10 = w2;
        a11 = dq;
        yj = yj_n;
        wq = wq_n;
        dq = dq_n;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  // column rotations of the block rows above the core (as in reg_chase_block)
  const int k0 = max(kstart, b0 + j + 1);
  if (len > 0 && j < nb && k0 <= kend) {
    cd* row = hb + j * ld - b0;
    cd xv = row[k0];
    int k = k0;
    for (; k + 4 <= kend + 1; k += 4) {
      cd yv[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) yv[m] = row[k + 1 + m];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        rot_cols(xv, yv[m], crot[k + m], srot[k + m]);
        row[k + m] = xv;
        xv = yv[m];
      }
    }
    for (; k <= kend; ++k) {
      cd yv = row[k + 1];
      rot_cols(xv, yv, crot[k], srot[k]);
      row[k] = xv;
      xv = yv;
    }
    row[kend + 1] = xv;
  }
}

// One matrix of a paired sweep.
struct PairSweep {
  cd* H;
  cd* hb;
  double* crot;
  cd* srot;
  cd shift;
  int l, ihi;
};

// qr_sweep_reg for two matrices (on[m]: matrix m takes a sweep), blocks of
// EIG_PQB rows.  Per matrix the phases and their order are those of
// qr_sweep_reg: store the block; rows above the last block, or the next
// block's new columns; load the next block; chase it while the other
// wavefronts apply the finished block's rotations to the rest of the window.
// The line work of both matrices shares one index space.
template <int EIG_THREADS>
__device__ __forceinline__ void qr_sweep_pair(const PairSweep (&S)[2], const bool (&on)[2], int n,
                                              unsigned long long* acc) {
  constexpr int RW = EIG_THREADS - 64;
  const int tid = threadIdx.x, lane = tid & 63;
  const int h = lane >> 5;  // the chase wavefront's half
  int kstart[2], kend[2], b0[2], b1[2];
  bool run[2];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    run[m] = on[m];
    kstart[m] = S[m].l;
    b0[m] = S[m].l;
    b1[m] = min(b0[m] + EIG_PQB - 1, S[m].ihi);
    kend[m] = (b1[m] == S[m].ihi) ? S[m].ihi - 1 : b1[m] - 2;
    if (run[m]) {
      const int nb = b1[m] - b0[m] + 1, ld = nb + 1;
      for (int e = tid; e < nb * nb; e += EIG_THREADS)
        if (e % nb >= e / nb - 2) S[m].hb[(e / nb) * ld + e % nb] = S[m].H[(long long)(b0[m] + e / nb) * n + b0[m] + e % nb];
    }
  }
  __syncthreads();
  unsigned long long t_c0 = acc ? wall_clock64() : 0;
  if (tid < 64) {
    const PairSweep& P = h ? S[1] : S[0];
    const int pb0 = h ? b0[1] : b0[0], pnb = (h ? b1[1] : b1[0]) - pb0 + 1;
    reg_chase_pair(h ? run[1] : run[0], P.hb, pnb, pb0, P.l, h ? kstart[1] : kstart[0], h ? kend[1] : kend[0],
                   P.shift, P.crot, P.srot, lane, pnb + 1);
  }
  __syncthreads();
  if (acc) acc[1] += wall_clock64() - t_c0;
  while (run[0] || run[1]) {
    bool last[2];
    int b0n[2], b1n[2], ksn[2], ken[2], cnt[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      last[m] = kend[m] == S[m].ihi - 1;
      b0n[m] = kend[m];
      b1n[m] = min(b0n[m] + EIG_PQB - 1, S[m].ihi);
      ksn[m] = kend[m] + 1;
      ken[m] = (b1n[m] == S[m].ihi) ? S[m].ihi - 1 : b1n[m] - 2;
      cnt[m] = !run[m] ? 0 : last[m] ? b0[m] - S[m].l : b1n[m] - b1[m];
      if (run[m]) {
        const int nb = b1[m] - b0[m] + 1, ld = nb + 1;
        for (int e = tid; e < nb * nb; e += EIG_THREADS)
          if (e % nb >= e / nb - 2) S[m].H[(long long)(b0[m] + e / nb) * n + b0[m] + e % nb] = S[m].hb[(e / nb) * ld + e % nb];
      }
    }
    // the last block's rows above (column rotations) or the next block's new
    // columns b1+1..b1n (row rotations)
    for (int e = tid; e < cnt[0] + cnt[1]; e += EIG_THREADS) {
      const int m = e >= cnt[0];
      const int i = m ? e - cnt[0] : e;
      const PairSweep& P = m ? S[1] : S[0];
      const int ks = m ? kstart[1] : kstart[0], ke = m ? kend[1] : kend[0];
      if (m ? last[1] : last[0])
        slide<false, EIG_SLIDE_PF>(P.H + (long long)(P.l + i) * n, 1, ks, ke + 1, P.crot, P.srot);
      else
        slide<true, EIG_SLIDE_PF>(P.H + (m ? b1[1] : b1[0]) + 1 + i, n, ks, ke + 1, P.crot, P.srot);
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      if (run[m] && last[m]) run[m] = false;
      if (run[m]) {
        const int nb = b1n[m] - b0n[m] + 1, ld = nb + 1;
        for (int e = tid; e < nb * nb; e += EIG_THREADS)
          if (e % nb >= e / nb - 2) S[m].hb[(e / nb) * ld + e % nb] = S[m].H[(long long)(b0n[m] + e / nb) * n + b0n[m] + e % nb];
      }
    }
    if (!(run[0] || run[1])) break;
    __syncthreads();
    t_c0 = acc ? wall_clock64() : 0;
    if (tid < 64) {
      const PairSweep& P = h ? S[1] : S[0];
      const int pb0 = h ? b0n[1] : b0n[0], pnb = (h ? b1n[1] : b1n[0]) - pb0 + 1;
      reg_chase_pair(h ? run[1] : run[0], P.hb, pnb, pb0, P.l, h ? ksn[1] : ksn[0], h ? ken[1] : ken[0], P.shift,
                     P.crot, P.srot, lane, pnb + 1);
    } else {
      // rest of the finished blocks: rows l..b0-1 (column rotations) and
      // columns b1n+1..ihi (row rotations)
      int na[2], tot[2];
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        na[m] = b0[m] - S[m].l;
        tot[m] = run[m] ? na[m] + S[m].ihi - b1n[m] : 0;
      }
      for (int e = tid - 64; e < tot[0] + tot[1]; e += RW) {
        const int m = e >= tot[0];
        const int i = m ? e - tot[0] : e;
        const PairSweep& P = m ? S[1] : S[0];
        const int ks = m ? kstart[1] : kstart[0], ke = m ? kend[1] : kend[0];
        const int nam = m ? na[1] : na[0];
        if (i < nam)
          slide<false, EIG_SLIDE_PF>(P.H + (long long)(P.l + i) * n, 1, ks, ke + 1, P.crot, P.srot);
        else
          slide<true, EIG_SLIDE_PF>(P.H + (m ? b1n[1] : b1n[0]) + 1 + (i - nam), n, ks, ke + 1, P.crot, P.srot);
      }
    }
    __syncthreads();
    if (acc) acc[1] += wall_clock64() - t_c0;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      if (run[m]) {
        kstart[m] = ksn[m];
        kend[m] = ken[m];
        b0[m] = b0n[m];
        b1[m] = b1n[m];
      }
    }
  }
}

// ------------------------------------------------------------ multishift QR
// A chain pass applies MS_NS shifts at once: MS_NS bulges two positions apart
// chased together (the small-bulge multishift QR of Braman, Byers and
// Mathias, with complex Givens bulges).  At step t bulge b is at
// k = l + t - 2b and active while l <= k < ihi; within a step the bulges act
// on disjoint row pairs and disjoint column pairs, and a row rotation
// commutes with a column rotation, so the pass produces the matrix of MS_NS
// single-shift sweeps applied one after the other (up to rounding).
//
// The chase runs through LDS windows of MS_W rows / columns on one
// wavefront: per step lane b computes bulge b's rotation, then lane j applies
// every bulge's row rotation to window column j and lane i every column
// rotation to window row i.  Outside the window:
//   * columns right of the window inside the active block (the next windows
//     load them) take the window's row rotations right after its chase;
//   * rows above the window are never read again in the pass, columns right
//     of the active block and Zt are never read by the chase: they take all
//     their rotations in one sweep per line at the end of the pass.
// Per pass each line is read and written once (twice for the columns inside
// the active block), instead of once per sweep.


__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Wilkinson shift: the eigenvalue of [[a, b], [c, d]] closest to d.
__device__ __forceinline__ cd wilkinson_shift(cd a, cd b, cd c, cd d) {
  const cd bc = b * c;
  if (bc.x == 0.0 && bc.y == 0.0) return d;
  const cd tr = 0.5 * (a - d);
  const cd disc = sqrtc(fmac(bc, tr, tr));
  const cd p1 = tr + disc, p2 = tr - disc;
  const cd den = abs2c(p1) >= abs2c(p2) ? p1 : p2;
  return (den.x == 0.0 && den.y == 0.0) ? d : d - divc(bc, den);
}

// Shifts of a chain pass (wavefront 0): the eigenvalues of the trailing
// MS_NS x MS_NS block of H, by single-shift QR of an LDS copy hs (lane j
// takes column j of a row rotation, lane i row i of a column rotation), or
// when `exceptional` the block's diagonal moved by 0.75 times its
// subdiagonal magnitudes (zlaqr0's exceptional shifts).  Without convergence
// the block's diagonal is used.  sh gets them ordered by decreasing distance
// from H[ihi][ihi], so the closest is chased last.
__device__ __forceinline__ void ms_shifts(const cd* H, int n, int ihi, bool exceptional, cd* hs, cd* sh, int lane) {
  constexpr int m = MS_NS, ld = MS_NS + 1;
  const int i0 = ihi - m + 1;
  if (exceptional) {
    if (lane < m) {
      const int i = i0 + lane;
      double sabs = cabs1(H[(long long)i * n + i - 1]);
      if (i >= 2) sabs += cabs1(H[(long long)(i - 1) * n + i - 2]);
      hs[lane * ld + lane] = H[(long long)i * n + i] + mkc(0.75 * sabs, 0.0);
    }
    wave_sync();
  } else {
    for (int e = lane; e < m * m; e += 64) {
      const int r = e / m, c = e % m;
      hs[r * ld + c] = (c >= r - 1) ? H[(long long)(i0 + r) * n + i0 + c] : mkc(0.0, 0.0);
    }
    wave_sync();
    const double smlnum = SAFMIN * ((double)m / ULP);
    int hi = m - 1, its = 0, tot = 0;
    while (hi >= 1 && tot < 30 * m) {
      int lk = 0;
      if (lane >= 1 && lane <= hi && negligible(hs, ld, lane, smlnum)) lk = lane;
      const int l = __builtin_amdgcn_readfirstlane(wave_max_int(lk));
      wave_sync();
      if (lane == l && l > 0) hs[l * ld + l - 1] = mkc(0.0, 0.0);
      if (l == hi) {
        --hi;
        its = 0;
        wave_sync();
        continue;
      }
      ++its;
      ++tot;
      cd shift;
      if (its % 10 == 0)
        shift = hs[hi * ld + hi] + mkc(0.75 * fabs(hs[hi * ld + hi - 1].x), 0.0);
      else
        shift = wilkinson_shift(hs[(hi - 1) * ld + hi - 1], hs[(hi - 1) * ld + hi], hs[hi * ld + hi - 1],
                                hs[hi * ld + hi]);
      cd x = hs[l * ld + l] - shift, y = hs[(l + 1) * ld + l];
      for (int k = l; k < hi; ++k) {
        if (k > l) {
          x = hs[k * ld + k - 1];
          y = hs[(k + 1) * ld + k - 1];
        }
        double c;
        cd s, r;
        givens(x, y, c, s, r);
        wave_sync();  // every lane has read x, y
        const int j = lane;
        if (k > l && j == k - 1) {
          hs[k * ld + j] = r;
          hs[(k + 1) * ld + j] = mkc(0.0, 0.0);
        } else if (j >= k && j <= hi) {
          cd a = hs[k * ld + j], b = hs[(k + 1) * ld + j];
          rot_rows(a, b, c, s);
          hs[k * ld + j] = a;
          hs[(k + 1) * ld + j] = b;
        }
        wave_sync();
        const int i = lane;
        if (i >= l && i <= min(k + 2, hi)) {
          cd a = hs[i * ld + k], b = hs[i * ld + k + 1];
          rot_cols(a, b, c, s);
          hs[i * ld + k] = a;
          hs[i * ld + k + 1] = b;
        }
        wave_sync();
      }
    }
  }
  if (lane == 0) {  // selection sort into sh, farthest from H[ihi][ihi] first
    const cd ref = H[(long long)ihi * n + ihi];
    for (int i = 0; i < m; ++i) sh[i] = hs[i * ld + i];
    for (int i = 0; i < m; ++i) {
      int best = i;
      double bd = abs2c(sh[i] - ref);
      for (int j = i + 1; j < m; ++j) {
        const double d = abs2c(sh[j] - ref);
        if (d > bd) {
          bd = d;
          best = j;
        }
      }
      const cd tmp = sh[i];
      sh[i] = sh[best];
      sh[best] = tmp;
    }
  }
  wave_sync();
}

__device__ __forceinline__ bool ms_active(int l, int ihi, int t, int b, int& k) {
  k = l + t - 2 * b;
  return k >= l && k < ihi;
}

// Chase steps [t0, te) inside the LDS window hb (rows / columns w0 .. w0+nb-1,
// row stride ld) on wavefront 0; the rotation of bulge b at step t goes to
// slot (t - t0) MS_NS + b of ps / pc.
__device__ __forceinline__ void ms_chase(cd* hb, int ld, int nb, int w0, int l, int ihi, int t0, int te, const cd* sh, cd* ps,
                         double* pc, int lane) {
#pragma unroll 1
  for (int t = t0; t < te; ++t) {
    const int sl = (t - t0) * MS_NS;
    if (lane < MS_NS) {
      int k;
      if (ms_active(l, ihi, t, lane, k)) {
        const int kk = k - w0;
        cd x, y;
        if (k == l) {
          x = hb[kk * ld + kk] - sh[lane];
          y = hb[(kk + 1) * ld + kk];
        } else {
          x = hb[kk * ld + kk - 1];
          y = hb[(kk + 1) * ld + kk - 1];
        }
        double c;
        cd s, r;
        givens(x, y, c, s, r);
        pc[sl + lane] = c;
        ps[sl + lane] = s;
        if (k > l) {  // the bulge this rotation removes
          hb[kk * ld + kk - 1] = r;
          hb[(kk + 1) * ld + kk - 1] = mkc(0.0, 0.0);
        }
      }
    }
    wave_sync();
    // row rotations: lane = window column; every bulge's operands are loaded
    // first (inactive ones from a valid dummy address), then rotated, then
    // stored, so the LDS round trips of the bulges overlap
    {
      cd xv[MS_NS], yv[MS_NS];
      int off[MS_NS];
      bool on[MS_NS];
#pragma unroll
      for (int b = 0; b < MS_NS; ++b) {
        int k;
        const bool act = ms_active(l, ihi, t, b, k);
        const int kk = k - w0;
        on[b] = act && lane >= kk && lane < nb;
        off[b] = on[b] ? kk * ld + lane : 0;
        xv[b] = hb[off[b]];
        yv[b] = hb[off[b] + ld];
      }
#pragma unroll
      for (int b = 0; b < MS_NS; ++b) rot_rows(xv[b], yv[b], pc[sl + b], ps[sl + b]);
#pragma unroll
      for (int b = 0; b < MS_NS; ++b)
        if (on[b]) {
          hb[off[b]] = xv[b];
          hb[off[b] + ld] = yv[b];
        }
    }
    wave_sync();
    // column rotations: lane = window row
    {
      cd xv[MS_NS], yv[MS_NS];
      int off[MS_NS];
      bool on[MS_NS];
#pragma unroll
      for (int b = 0; b < MS_NS; ++b) {
        int k;
        const bool act = ms_active(l, ihi, t, b, k);
        const int kk = k - w0;
        on[b] = act && lane <= kk + 2 && lane < nb;
        off[b] = on[b] ? lane * ld + kk : 0;
        xv[b] = hb[off[b]];
        yv[b] = hb[off[b] + 1];
      }
#pragma unroll
      for (int b = 0; b < MS_NS; ++b) rot_cols(xv[b], yv[b], pc[sl + b], ps[sl + b]);
#pragma unroll
      for (int b = 0; b < MS_NS; ++b)
        if (on[b]) {
          hb[off[b]] = xv[b];
          hb[off[b] + 1] = yv[b];
        }
    }
    wave_sync();
  }
}

// One line v[p * step] (p = matrix index along the line; only p in [lo, hi]
// is loaded or stored) taking the rotations of a stretch of chain steps:
// ROWS = row rotations (a column of H), else column rotations (a row of H)
// or their conjugates (a column of Zt).  Rotation (t, b) acts on the pair
// (l + t - 2b, l + t - 2b + 1); the MS_R elements around the chain stay in
// registers (r[i] = element l + t - MS_LAG + i at step t), one element
// entering and one leaving per step.
struct MsLine {
  cd r[MS_R];
  cd pf[MS_PF];
};

__device__ __forceinline__ cd ms_ld(const cd* v, long long step, int e, int lo, int hi) {
  return (e >= lo && e <= hi) ? v[(long long)e * step] : mkc(0.0, 0.0);
}

__device__ __forceinline__ void ms_line_begin(MsLine& L, const cd* v, long long step, int lo, int hi, int l,
                                              int ts) {
  const int b0 = l + ts - MS_LAG;
#pragma unroll
  for (int i = 0; i < MS_R - 1; ++i) L.r[i] = ms_ld(v, step, b0 + i, lo, hi);
  L.r[MS_R - 1] = mkc(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < MS_PF; ++i) L.pf[i] = ms_ld(v, step, l + ts + 1 + i, lo, hi);
}

// steps [t0, t1); parameters of step t at slot (t - pbase) MS_NS + b of ps / pc
template <bool ROWS>
__device__ __forceinline__ void ms_line_steps(MsLine& L, cd* v, long long step, int lo, int hi, int l, int ihi,
                                              int t0, int t1, const cd* ps, const double* pc, int pbase) {
#pragma unroll 1
  for (int t = t0; t < t1; ++t) {
    L.r[MS_R - 1] = L.pf[0];
#pragma unroll
    for (int i = 0; i + 1 < MS_PF; ++i) L.pf[i] = L.pf[i + 1];
    L.pf[MS_PF - 1] = ms_ld(v, step, l + t + 1 + MS_PF, lo, hi);
    const int sl = (t - pbase) * MS_NS;
#pragma unroll
    for (int b = 0; b < MS_NS; ++b) {
      const int k = l + t - 2 * b;
      if (k >= l && k < ihi) {
        const double c = pc[sl + b];
        const cd s = ps[sl + b];
        if (ROWS)
          rot_rows(L.r[MS_LAG - 2 * b], L.r[MS_LAG - 2 * b + 1], c, s);
        else
          rot_cols(L.r[MS_LAG - 2 * b], L.r[MS_LAG - 2 * b + 1], c, s);
      }
    }
    const int eo = l + t - MS_LAG;
    if (eo >= lo && eo <= hi) v[(long long)eo * step] = L.r[0];
#pragma unroll
    for (int i = 0; i + 1 < MS_R; ++i) L.r[i] = L.r[i + 1];
  }
}

__device__ __forceinline__ void ms_line_end(const MsLine& L, cd* v, long long step, int lo, int hi, int l, int te) {
  const int b0 = l + te - MS_LAG;
#pragma unroll
  for (int i = 0; i < MS_R - 1; ++i) {
    const int e = b0 + i;
    if (e >= lo && e <= hi) v[(long long)e * step] = L.r[i];
  }
}

// End of a chain pass: every row of H up to ihi takes the column rotations
// from the step it left the windows (tleave), every column of Zt the
// conjugate column rotations and every column right of the active block the
// row rotations of the whole pass (the log gs / gc, slot t MS_NS + b), staged
// MS_CH steps at a time in LDS.  One loop per rotation kind (ROWS), so a
// thread's line state is one register window.
template <int EIG_THREADS, bool ROWS>
__device__ __forceinline__ void ms_deferred_kind(cd* H, cd* Zt, int n, int l, int ihi, int T, const int* tleave,
                                                 const cd* gs, const double* gc, cd* stage) {
  cd* st_s = stage;
  double* st_c = reinterpret_cast<double*>(stage + MS_CH * MS_NS);
  // ROWS: columns ihi+1..n-1 of H; else rows 0..ihi of H, then the columns of Zt
  const int nlines = ROWS ? n - 1 - ihi : ihi + 1 + n;
#pragma unroll 1
  for (int base = 0; base < nlines; base += EIG_THREADS) {
    const int e = base + threadIdx.x;
    cd* v = H;
    long long step = ROWS ? n : 1;
    int ts = T;
    if (ROWS) {
      if (e < nlines) {
        v = H + ihi + 1 + e;
        ts = 0;
      }
    } else if (e <= ihi) {
      v = H + (long long)e * n;
      ts = tleave[e];
    } else if (e < nlines) {
      v = Zt + (e - ihi - 1);
      step = n;
      ts = 0;
    }
    MsLine L;
    bool started = false;
#pragma unroll 1
    for (int c0 = 0; c0 < T; c0 += MS_CH) {
      const int c1 = min(c0 + MS_CH, T);
      __syncthreads();  // the previous chunk is consumed
#pragma unroll 1
      for (int idx = threadIdx.x; idx < (c1 - c0) * MS_NS; idx += EIG_THREADS) {
        st_s[idx] = gs[(long long)c0 * MS_NS + idx];
        st_c[idx] = gc[(long long)c0 * MS_NS + idx];
      }
      __syncthreads();
      if (ts < c1) {
        if (!started) {
          ms_line_begin(L, v, step, l, ihi, l, ts);
          started = true;
        }
        ms_line_steps<ROWS>(L, v, step, l, ihi, l, ihi, max(c0, ts), c1, st_s, st_c, c0);
      }
    }
    if (started) ms_line_end(L, v, step, l, ihi, l, T);
  }
}

template <int EIG_THREADS>
__device__ __forceinline__ void ms_deferred(cd* H, cd* Zt, int n, int l, int ihi, int T, const int* tleave,
                                            const cd* gs, const double* gc, cd* stage) {
  ms_deferred_kind<EIG_THREADS, false>(H, Zt, n, l, ihi, T, tleave, gs, gc, stage);
  if (ihi + 1 < n) ms_deferred_kind<EIG_THREADS, true>(H, Zt, n, l, ihi, T, tleave, gs, gc, stage);
  __syncthreads();
}

// One chain pass over the active block [l, ihi] with the shifts sh (LDS).
// hb: the LDS window (MS_W x (MS_W + 1)), ps / pc: the window's rotation
// slots (LDS), tleave: n ints (LDS), gs / gc: the pass's rotation log
// (global, (ihi - l + MS_LAG) MS_NS slots).
template <int EIG_THREADS>
__device__ __forceinline__ void ms_chain_pass(cd* H, cd* Zt, int n, int l, int ihi, const cd* sh, cd* hb, cd* ps, double* pc,
                              int* tleave, cd* gs, double* gc, unsigned long long* acc) {
  const int tid = threadIdx.x;
  const int T = (ihi - l) + MS_LAG;  // the trailing bulge's last rotation is at step T - 1
  for (int i = tid; i <= ihi; i += EIG_THREADS) tleave[i] = i < l ? 0 : T;
  const unsigned long long t_p0 = acc ? wall_clock64() : 0;
  int t = 0, w0 = l;
  while (true) {
    const int w1 = min(w0 + MS_W - 1, ihi);
    const int nb = w1 - w0 + 1, ld = nb + 1;
    // steps [t, te) whose rotations stay inside the window: each active
    // bulge reads column k - 1 (or starts at k = l) and writes row k + 2 (or
    // is at its last rotation k = ihi - 1)
    int te = t;
    while (te < T && te - t < MS_W) {
      bool ok = true;
#pragma unroll
      for (int b = 0; b < MS_NS; ++b) {
        int k;
        if (!ms_active(l, ihi, te, b, k)) continue;
        if (k == l ? k < w0 : k - 1 < w0) ok = false;
        if (!(k + 2 <= w1 || (k + 1 == ihi && ihi <= w1))) ok = false;
      }
      if (!ok) break;
      ++te;
    }
    __syncthreads();  // tleave, the previous window's column updates
    for (int e = tid; e < nb * nb; e += EIG_THREADS) {
      const int r = e / nb, c = e % nb;
      if (c >= r - 2) hb[r * ld + c] = H[(long long)(w0 + r) * n + w0 + c];
    }
    __syncthreads();
    const unsigned long long t_c0 = acc ? wall_clock64() : 0;
    if (tid < 64 && te > t) ms_chase(hb, ld, nb, w0, l, ihi, t, te, sh, ps, pc, tid);
    __syncthreads();
    if (acc) acc[1] += wall_clock64() - t_c0;
    if (te == t) break;  // cannot happen (the chain always fits a window of MS_W >= 2 MS_NS + 3)
    for (int e = tid; e < nb * nb; e += EIG_THREADS) {
      const int r = e / nb, c = e % nb;
      if (c >= r - 2) H[(long long)(w0 + r) * n + w0 + c] = hb[r * ld + c];
    }
    // columns w1+1..ihi (rows w0..w1) take this window's row rotations; the
    // window's slots go to the pass log
    for (int j = w1 + 1 + tid; j <= ihi; j += EIG_THREADS) {
      MsLine L;
      ms_line_begin(L, H + j, n, w0, w1, l, t);
      ms_line_steps<true>(L, H + j, n, w0, w1, l, ihi, t, te, ps, pc, t);
      ms_line_end(L, H + j, n, w0, w1, l, te);
    }
    for (int idx = tid; idx < (te - t) * MS_NS; idx += EIG_THREADS) {
      gs[(long long)t * MS_NS + idx] = ps[idx];
      gc[(long long)t * MS_NS + idx] = pc[idx];
    }
    t = te;
    if (t >= T) break;
    // next window: from the trailing bulge's input column, or from l while
    // bulges are still entering; rows above it have left the windows
    const int w0n = (t <= MS_LAG) ? l : max(l, l + t - MS_LAG - 1);
    for (int i = w0 + tid; i < w0n; i += EIG_THREADS) tleave[i] = t;
    w0 = w0n;
  }
  __syncthreads();
  const unsigned long long t_d0 = acc ? wall_clock64() : 0;
  if (acc) acc[2] += t_d0 - t_p0;
  ms_deferred<EIG_THREADS>(H, Zt, n, l, ihi, T, tleave, gs, gc, hb);
  if (acc) acc[4] += wall_clock64() - t_d0;
}

// Stages 1-3.  A (n x n) becomes T (upper triangle); Zt gets the transposed
// Schur vectors; X the eigenvectors of T (upper triangular, unit diagonal);
// lam the eigenvalues.  STAGES = 7: all three in one workgroup with H in LDS
// (n <= KHX_EIG_LDS_N); otherwise one stage (1 Hessenberg, 2 Schur form,
// 4 eigenvectors) per launch with H in global memory, so that each stage's
// kernel gets the registers it needs (the chain passes of the Schur stage
// would spill inside a kernel that also holds the other stages).
template <int EIG_THREADS, int STAGES, bool CHAIN = false>
__device__ __forceinline__ void eig_body(int n, cd* A, long long sA, cd* Ztg, cd* Xg, long long sW, cd* lam,
                                         long long sL, int* info, unsigned long long* tphase, int flags) {
  constexpr int EIG_WAVES = EIG_THREADS / 64;
  extern __shared__ cd sm[];
  // region 1: u, vh, srot, crot (Hessenberg, single-shift sweeps,
  // eigenvectors) or, in a chain pass, the window's rotation slots ms_ps /
  // ms_pc; then wv (small n), part (Hessenberg sums / T, the QR windows, the
  // staged rotation parameters), and H itself (small n) or the chain pass's
  // per-row step tables and shifts.
  cd* u = sm;                                           // n     Householder vector / solved row
  cd* vh = u + n;                                       // n     H u
  cd* srot = vh + n;                                    // n     rotation sines
  double* crot = reinterpret_cast<double*>(srot + n);   // n     cosines
  cd* ms_ps = sm;                                       // MS_SLOTS sines of a chain window
  double* ms_pc = reinterpret_cast<double*>(sm + MS_SLOTS);  // MS_SLOTS cosines
  cd* wv = sm + eig_r1_elems(n, CHAIN);                 // 2n    u^H H and u^T Zt (n <= KHX_EIG_LDS_N)
  cd* part = wv + eig_wv_elems(n);                      // partial sums, QR windows, staged parameters
  cd* hl = part + eig_part_elems(n, EIG_WAVES, CHAIN);  // n*n LDS copy of H (small n)
  int* ms_tl = reinterpret_cast<int*>(hl);              // n     step each row leaves the chain windows
  cd* ms_sh = hl + (n + 3) / 4;                         // MS_NS shifts of a chain pass
  __shared__ double red[EIG_WAVES];
  __shared__ int redi[EIG_WAVES];
  __shared__ cd s_xy[2][2];  // double-buffered inputs of the next rotation

  const long long item = blockIdx.x;
  cd* Hg = A + item * sA;
  cd* Zt = Ztg + item * sW;
  cd* X = Xg + item * sW;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const double smlnum = SAFMIN * ((double)n / ULP);
  constexpr bool in_lds = STAGES == 7;
  cd* H = in_lds ? hl : Hg;
  // optional phase timestamps (wall clock): start, Hessenberg, Schur, eigenvectors, sweeps
  unsigned long long* tp = tphase ? tphase + item * 10 : nullptr;
  unsigned long long acc_t[5] = {0, 0, 0, 0, 0};  // scan, chase, block rest, log, flush

  if (STAGES & 1) {
  if (tp && tid == 0) tp[0] = wall_clock64();
  for (int i = wid; i < n; i += EIG_WAVES)
    for (int j = lane; j < n; j += 64) {
      Zt[i * n + j] = mkc(i == j ? 1.0 : 0.0, 0.0);
      if (in_lds) hl[i * n + j] = Hg[(long long)i * n + j];
    }
  __syncthreads();

  // ------------------------------------------------------------ 1. Hessenberg
  if (!in_lds)
  {
    if (flags & 8)
      hessenberg_blocked<EIG_THREADS, true>(H, Zt, n, X, X + (long long)n * HB_NB, X + 2ll * n * HB_NB, u, vh, part,
                                            red);
    else
      hessenberg_blocked<EIG_THREADS, false>(H, Zt, n, X, X + (long long)n * HB_NB, X + 2ll * n * HB_NB, u, vh, part,
                                             red);
  }
  for (int k = 0; in_lds && k + 2 < n; ++k) {
    const int L = n - k - 1;  // reflector acts on rows/cols k+1..n-1
    double part2 = 0.0;
    for (int i = k + 2 + tid; i < n; i += EIG_THREADS) part2 += abs2c(H[i * n + k]);
    const double rest2 = block_sum<EIG_THREADS>(part2, red);
    if (rest2 == 0.0) continue;  // uniform: every thread sees the same sum
    const cd alpha = H[(k + 1) * n + k];
    const double aa = absc(alpha);
    const double xn = sqrt(aa * aa + rest2);
    const cd ph = aa != 0.0 ? (1.0 / aa) * alpha : mkc(1.0, 0.0);
    const double sc = 1.0 / sqrt(xn * (xn + aa));
    for (int i = tid; i < L; i += EIG_THREADS) {
      cd x = H[(k + 1 + i) * n + k];
      if (i == 0) x = x + xn * ph;
      u[i] = sc * x;
    }
    __syncthreads();
    // partial sums: H columns j > k (conj(u)^T H) and Zt columns (u^T Zt)
    for (int j = lane; j < 2 * n; j += 64) {
      const bool hz = j < n;
      const int col = hz ? j : j - n;
      if (hz && col <= k) continue;
      cd acc = mkc(0.0, 0.0);
      if (hz) {
#pragma unroll 8
        for (int i = wid; i < L; i += EIG_WAVES) acc = fmac_cj(acc, u[i], H[(k + 1 + i) * n + col]);
      } else {
#pragma unroll 8
        for (int i = wid; i < L; i += EIG_WAVES) acc = fmac(acc, u[i], Zt[(k + 1 + i) * n + col]);
      }
      part[wid * 2 * n + j] = acc;
    }
    __syncthreads();
    for (int j = tid; j < 2 * n; j += EIG_THREADS) {
      if (j < n && j <= k) continue;
      cd acc = part[j];
#pragma unroll
      for (int w = 1; w < EIG_WAVES; ++w) acc = acc + part[w * 2 * n + j];
      wv[j] = acc;
    }
    __syncthreads();
    // H := P H (rows k+1..), Zt := conj(P) Zt (Z := Z P)
    for (int i = wid; i < L; i += EIG_WAVES) {
      const cd ui = u[i], cui = conjc(u[i]);
      cd* row = H + (k + 1 + i) * n;
      for (int j = k + 1 + lane; j < n; j += 64) row[j] = fmsc(row[j], ui, wv[j]);
      if (lane == 0) row[k] = (i == 0) ? -(xn * ph) : mkc(0.0, 0.0);
      cd* zr = Zt + (k + 1 + i) * n;
      for (int j = lane; j < n; j += 64) zr[j] = fmsc(zr[j], cui, wv[n + j]);
    }
    __syncthreads();
    // H := H P (columns k+1..): v_i = sum_j H[i][k+1+j] u_j, one thread per row
    for (int i = tid; i < n; i += EIG_THREADS) {
      const cd* hr = H + i * n + k + 1;
      cd a0 = mkc(0.0, 0.0), a1 = mkc(0.0, 0.0);
      int j = 0;
#pragma unroll 4
      for (; j + 1 < L; j += 2) {
        a0 = fmac(a0, hr[j], u[j]);
        a1 = fmac(a1, hr[j + 1], u[j + 1]);
      }
      if (j < L) a0 = fmac(a0, hr[j], u[j]);
      vh[i] = a0 + a1;
    }
    __syncthreads();
    for (int i = wid; i < n; i += EIG_WAVES) {
      cd* hr = H + i * n + k + 1;
      const cd a = vh[i];
      for (int j = lane; j < L; j += 64) hr[j] = fmsc(hr[j], a, conjc(u[j]));
    }
    __syncthreads();
  }

  if (tp && tid == 0) tp[1] = wall_clock64();
  }  // stage 1
  // ------------------------------------------------------------ 2. Schur form
  if (STAGES & 2) {
  int ihi = n - 1, its = 0, itot = 0, fail = 0;
  // rotation log of the sweeps not yet applied outside their windows (X is
  // free until the eigenvector stage)
  __shared__ int s_lo[EIG_KLOG], s_hi[EIG_KLOG];
  int nlog = 0;
  cd* log_s = X;
  double* log_c = reinterpret_cast<double*>(X + (long long)EIG_KLOG * n);
  // iteration limit; bits 8.. of flags lower it (KHX_EIG_ITMAX: testing the
  // non-convergence status without a matrix that fails to converge)
  const int itmax = (flags >> 8) > 0 ? (flags >> 8) : 30 * max(10, n);
  // chain passes (multishift) in the rest of the log area of X
  cd* ms_gs = X + 10ll * n;
  double* ms_gc = reinterpret_cast<double*>(ms_gs + (long long)(n + MS_LAG) * MS_NS);
  constexpr bool multishift = !in_lds && CHAIN;
  int npass = 0;  // chain passes among the itot iterations
  while (ihi >= 1) {
    const unsigned long long t_s0 = tp ? wall_clock64() : 0;
    int lk = 0;
    for (int k = 1 + tid; k <= ihi; k += EIG_THREADS) {
      if (negligible(H, n, k, smlnum)) {
        H[k * n + k - 1] = mkc(0.0, 0.0);
        lk = max(lk, k);
      }
    }
    const int l = block_max_int<EIG_THREADS>(lk, redi);
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
    if (multishift && ihi - l + 1 >= MS_MIN_WL) {
      unsigned long long t1 = tp ? wall_clock64() : 0;
      if (tp) acc_t[0] += t1 - t_s0;
      if (nlog > 0) {  // the pass updates every line directly: logged sweeps go first
        flush_log<EIG_THREADS>(H, Zt, n, nlog, s_lo, s_hi, log_c, log_s, part);
        nlog = 0;
      }
      unsigned long long t2 = tp ? wall_clock64() : 0;
      if (tp) acc_t[4] += t2 - t1;
      if (tid < 64) ms_shifts(H, n, ihi, its % 6 == 0, part, ms_sh, lane);
      __syncthreads();
      if (tp) acc_t[3] += wall_clock64() - t2;
      ms_chain_pass<EIG_THREADS>(H, Zt, n, l, ihi, ms_sh, part, ms_ps, ms_pc, ms_tl, ms_gs, ms_gc,
                                 tp ? acc_t : nullptr);
      ++npass;
      continue;
    }
    cd shift;
    if (its % 20 == 10) {  // exceptional shifts break cycles
      shift = H[l * n + l] + mkc(0.75 * fabs(H[(l + 1) * n + l].x), 0.0);
    } else if (its % 20 == 0) {
      shift = H[ihi * n + ihi] + mkc(0.75 * fabs(H[ihi * n + ihi - 1].x), 0.0);
    } else {  // Wilkinson: eigenvalue of the trailing 2x2 closest to H[ihi][ihi]
      shift = wilkinson_shift(H[(ihi - 1) * n + ihi - 1], H[(ihi - 1) * n + ihi], H[ihi * n + ihi - 1],
                              H[ihi * n + ihi]);
    }
    if (!in_lds) {
      unsigned long long t1 = tp ? wall_clock64() : 0;
      if (tp) acc_t[0] += t1 - t_s0;
      // rows l.. of H must be current before the window reaches them
      if (nlog > 0 && l < s_lo[nlog - 1]) {
        flush_log<EIG_THREADS>(H, Zt, n, nlog, s_lo, s_hi, log_c, log_s, part);
        nlog = 0;
      }
      unsigned long long t2 = tp ? wall_clock64() : 0;
      if (tp) acc_t[4] += t2 - t1;
      if (flags & 4)
        qr_sweep_reg<EIG_THREADS>(H, n, l, ihi, shift, part, crot, srot, tp ? acc_t : nullptr);
      else
        qr_sweep_blocked<EIG_THREADS>(H, Zt, n, l, ihi, shift, part, crot, srot, &s_xy[0][0], (flags & 1) != 0,
                                      tp ? acc_t : nullptr);
      unsigned long long t3 = tp ? wall_clock64() : 0;
      if (tp) acc_t[2] += t3 - t2;
      for (int k = l + tid; k < ihi; k += EIG_THREADS) {
        log_c[nlog * n + k] = crot[k];
        log_s[nlog * n + k] = srot[k];
      }
      if (tid == 0) {
        s_lo[nlog] = l;
        s_hi[nlog] = ihi;
      }
      __syncthreads();
      unsigned long long t4 = tp ? wall_clock64() : 0;
      if (tp) acc_t[3] += t4 - t3;
      if (++nlog == EIG_KLOG) {
        flush_log<EIG_THREADS>(H, Zt, n, nlog, s_lo, s_hi, log_c, log_s, part);
        nlog = 0;
      }
      if (tp) acc_t[4] += wall_clock64() - t4;
      continue;
    }
    if (flags & 4) {
      // register chase over the whole window (at most 64 columns: one block)
      if (tid < 64) reg_chase_block(H + l * n + l, ihi - l + 1, l, l, ihi, l, ihi - 1, shift, crot, srot, tid, n);
      __syncthreads();
    } else {
    cd x = H[l * n + l] - shift;
    cd y = H[(l + 1) * n + l];
    // bulge chase restricted to the window [l, ihi]
    for (int k = l; k < ihi; ++k) {
      cd* nxt = s_xy[(k + 1) & 1];
      if (k > l) {
        x = s_xy[k & 1][0];
        y = s_xy[k & 1][1];
      }
      double c;
      cd s, r;
      givens(x, y, c, s, r);
      if (tid == 0) {
        crot[k] = c;
        srot[k] = s;
      }
      const int nrow = ihi - k - 1;                         // row part: columns k+2..ihi
      const int ncol = (k - l) + ((k + 2 <= ihi) ? 1 : 0);  // col part: rows l..k-1 (+ row k+2)
      const int total = 1 + nrow + ncol;
      for (int e = tid; e < total; e += EIG_THREADS) {
        if (e == 0) {
          if (k > l) {
            H[k * n + k - 1] = r;
            H[(k + 1) * n + k - 1] = mkc(0.0, 0.0);
          }
          cd* p0 = H + k * n + k;
          cd* p1 = H + (k + 1) * n + k;
          cd b00 = p0[0], b01 = p0[1], b10 = p1[0], b11 = p1[1];
          rot_rows(b00, b10, c, s);
          rot_rows(b01, b11, c, s);
          rot_cols(b00, b01, c, s);
          rot_cols(b10, b11, c, s);
          p0[0] = b00;
          p0[1] = b01;
          p1[0] = b10;
          p1[1] = b11;
          nxt[0] = b10;
        } else if (e <= nrow) {
          const int j = k + 1 + e;  // k+2 .. ihi
          cd xk = H[k * n + j], yk = H[(k + 1) * n + j];
          rot_rows(xk, yk, c, s);
          H[k * n + j] = xk;
          H[(k + 1) * n + j] = yk;
        } else {
          const int q = e - nrow - 1;
          const int i = (q < k - l) ? l + q : k + 2;
          cd* pr = H + i * n + k;
          cd xk = pr[0], yk = pr[1];
          rot_cols(xk, yk, c, s);
          pr[0] = xk;
          pr[1] = yk;
          if (i == k + 2) nxt[1] = xk;
        }
      }
      __syncthreads();
    }
    }
    // deferred: rows above the window, columns right of it, and Z
    const int nabove = l, nright = n - 1 - ihi;
    const int tot = nabove + nright + n;
    for (int e = tid; e < tot; e += EIG_THREADS) {
      if (e < nabove)  // row e of H, columns l..ihi
        slide<false>(H + e * n, 1, l, ihi, crot, srot);
      else if (e < nabove + nright)  // column j of H, rows l..ihi
        slide<true>(H + ihi + 1 + (e - nabove), n, l, ihi, crot, srot);
      else  // column i of Zt (row i of Z), rows l..ihi
        slide<false>(Zt + (e - nabove - nright), n, l, ihi, crot, srot);
    }
    __syncthreads();
  }
  if (nlog > 0) flush_log<EIG_THREADS>(H, Zt, n, nlog, s_lo, s_hi, log_c, log_s, part);
  if (tid == 0 && info) info[item] = fail;
  if (tp && tid == 0) {
    tp[2] = wall_clock64();
    tp[4] = (unsigned long long)itot | ((unsigned long long)npass << 32);
    tp[5] = acc_t[0];
    tp[6] = acc_t[1];
    tp[7] = acc_t[2] - acc_t[1];  // block load / store / in-window slides
    tp[8] = acc_t[3];
    tp[9] = acc_t[4];
  }

  cd* lm = lam + item * sL;
  for (int i = tid; i < n; i += EIG_THREADS) lm[i] = H[i * n + i];
  }  // stage 2

  // ------------------------------------------------------------ 3. eigenvectors of T
  if (STAGES & 4) {
  if (!in_lds && !(flags & 2)) {  // one thread per eigenvector (flags bit 1: the row-by-row sweep below)
    eigvec_columns<EIG_THREADS>(H, X, n, smlnum);
    if (tp && tid == 0) tp[3] = wall_clock64();
  } else {
  // X[i][k] = -T[i][k] (i < k), 1 (i == k), 0 (i > k)
  for (int i = wid; i < n; i += EIG_WAVES)
    for (int k = lane; k < n; k += 64) {
      const cd t = H[i * n + k];
      X[(long long)i * n + k] = (k > i) ? -t : mkc(k == i ? 1.0 : 0.0, 0.0);
    }
  __syncthreads();
  cd* xr = u;
  for (int j = n - 1; j >= 0; --j) {
    const cd tjj = H[j * n + j];
    for (int k = j + 1 + tid; k < n; k += EIG_THREADS) {
      const cd lk = H[k * n + k];
      cd d = tjj - lk;
      const double smin = fmax(ULP * cabs1(lk), smlnum);
      if (cabs1(d) < smin) d = mkc(smin, 0.0);
      const cd v = divc(X[(long long)j * n + k], d);
      X[(long long)j * n + k] = v;
      xr[k] = v;
    }
    __syncthreads();
    for (int i = wid; i < j; i += EIG_WAVES) {
      const cd tij = H[i * n + j];
      cd* row = X + (long long)i * n;
      for (int k = j + 1 + lane; k < n; k += 64) row[k] = fmsc(row[k], tij, xr[k]);
    }
    __syncthreads();
  }
  if (in_lds)  // the Schur form goes back to A (callers may inspect it)
    for (int i = wid; i < n; i += EIG_WAVES)
      for (int j = lane; j < n; j += 64) Hg[(long long)i * n + j] = hl[i * n + j];
  if (tp && tid == 0) tp[3] = wall_clock64();
  }
  }  // stage 3
}

// The Schur stage (single-shift sweeps, deferred rotation logs) of
// matrices 2 blockIdx.x and 2 blockIdx.x + 1 in one workgroup: both run the
// steps of eig_body's stage 2 -- deflation scan, shift, sweep, log, deferred
// pass -- with their sweeps taken together by qr_sweep_pair.  A matrix that
// has converged (or an absent second matrix of an odd batch) leaves the
// other to sweep alone.  Results are those of eig_body, bit for bit.
template <int EIG_THREADS>
__device__ __forceinline__ void eig_qr_pair(int n, cd* A, long long sA, cd* Ztg, cd* Xg, long long sW, cd* lam,
                                            long long sL, int* info, int batch, unsigned long long* tphase,
                                            int flags) {
  constexpr int EIG_WAVES = EIG_THREADS / 64, HT = EIG_THREADS / 2;
  extern __shared__ cd sm[];
  __shared__ int s_lo[2][EIG_KLOG], s_hi[2][EIG_KLOG];
  __shared__ int redi[EIG_WAVES];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const double smlnum = SAFMIN * ((double)n / ULP);
  const int itmax = (flags >> 8) > 0 ? (flags >> 8) : 30 * max(10, n);
  const int per = eig_pair_elems(n);
  PairSweep S[2];
  cd* Zt[2];
  cd* log_s[2];
  double* log_c[2];
  long long item[2];
  int ihi[2], its[2], itot[2], fail[2], nlog[2];
  bool live[2];
  unsigned long long acc_t[5] = {0, 0, 0, 0, 0};  // scan, chase, block rest, log, flush (the pair's wall time)
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    item[m] = 2ll * blockIdx.x + m;
    live[m] = item[m] < batch;
    const long long it = live[m] ? item[m] : 2ll * blockIdx.x;
    S[m].H = A + it * sA;
    Zt[m] = Ztg + it * sW;
    log_s[m] = Xg + it * sW;
    log_c[m] = reinterpret_cast<double*>(log_s[m] + (long long)EIG_KLOG * n);
    S[m].srot = sm + m * per;
    S[m].crot = reinterpret_cast<double*>(S[m].srot + n);
    S[m].hb = S[m].srot + n + (n + 1) / 2;
    S[m].shift = mkc(0.0, 0.0);
    S[m].l = 0;
    S[m].ihi = n - 1;
    ihi[m] = n - 1;
    its[m] = itot[m] = fail[m] = nlog[m] = 0;
  }
  while (live[0] || live[1]) {
    bool srch[2] = {live[0], live[1]}, sw[2] = {false, false};
    int l[2] = {0, 0};
    const unsigned long long t_s0 = tphase ? wall_clock64() : 0;
    while (srch[0] || srch[1]) {
      // deflation scans: the first half of the workgroup scans matrix 0, the second matrix 1
      const int m = tid >= HT;
      const int hi = m ? ihi[1] : ihi[0];
      cd* Hm = m ? S[1].H : S[0].H;
      int lk = 0;
      if (m ? srch[1] : srch[0]) {
        for (int k = 1 + tid - m * HT; k <= hi; k += HT) {
          if (negligible(Hm, n, k, smlnum)) {
            Hm[k * n + k - 1] = mkc(0.0, 0.0);
            lk = max(lk, k);
          }
        }
      }
      lk = wave_max_int(lk);
      if (lane == 0) redi[wid] = lk;
      __syncthreads();
      int lm[2] = {redi[0], redi[EIG_WAVES / 2]};
#pragma unroll
      for (int w = 1; w < EIG_WAVES / 2; ++w) {
        lm[0] = max(lm[0], redi[w]);
        lm[1] = max(lm[1], redi[EIG_WAVES / 2 + w]);
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (!srch[q]) continue;
        l[q] = __builtin_amdgcn_readfirstlane(lm[q]);
        if (l[q] == ihi[q]) {
          --ihi[q];
          its[q] = 0;
          if (ihi[q] < 1) srch[q] = live[q] = false;
          continue;
        }
        if (++itot[q] > itmax) {
          fail[q] = ihi[q] + 1;
          srch[q] = live[q] = false;
          continue;
        }
        ++its[q];
        const cd* H = S[q].H;
        const int lq = l[q], hq = ihi[q];
        if (its[q] % 20 == 10) {
          S[q].shift = H[lq * n + lq] + mkc(0.75 * fabs(H[(lq + 1) * n + lq].x), 0.0);
        } else if (its[q] % 20 == 0) {
          S[q].shift = H[hq * n + hq] + mkc(0.75 * fabs(H[hq * n + hq - 1].x), 0.0);
        } else {
          S[q].shift = wilkinson_shift(H[(hq - 1) * n + hq - 1], H[(hq - 1) * n + hq], H[hq * n + hq - 1],
                                       H[hq * n + hq]);
        }
        S[q].l = lq;
        S[q].ihi = hq;
        sw[q] = true;
        srch[q] = false;
      }
    }
    unsigned long long t1 = tphase ? wall_clock64() : 0;
    if (tphase) acc_t[0] += t1 - t_s0;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (!sw[q] && item[q] < batch && !live[q] && nlog[q] >= 0) {  // converged or failed just now
        if (nlog[q] > 0) flush_log<EIG_THREADS>(S[q].H, Zt[q], n, nlog[q], s_lo[q], s_hi[q], log_c[q], log_s[q], S[q].hb);
        nlog[q] = -1;
        if (tid == 0 && info) info[item[q]] = fail[q];
        if (tphase && tid == 0) {
          unsigned long long* tp = tphase + item[q] * 10;
          tp[2] = wall_clock64();
          tp[4] = (unsigned long long)itot[q];
          tp[5] = acc_t[0];
          tp[6] = acc_t[1];
          tp[7] = acc_t[2] - acc_t[1];
          tp[8] = acc_t[3];
          tp[9] = acc_t[4];
        }
        cd* lmq = lam + item[q] * sL;
        for (int i = tid; i < n; i += EIG_THREADS) lmq[i] = S[q].H[i * n + i];
      }
      // rows l.. of H must be current before the window reaches them
      if (sw[q] && nlog[q] > 0 && l[q] < s_lo[q][nlog[q] - 1]) {
        flush_log<EIG_THREADS>(S[q].H, Zt[q], n, nlog[q], s_lo[q], s_hi[q], log_c[q], log_s[q], S[q].hb);
        nlog[q] = 0;
      }
    }
    if (!(sw[0] || sw[1])) continue;
    unsigned long long t2 = tphase ? wall_clock64() : 0;
    if (tphase) acc_t[4] += t2 - t1;
    qr_sweep_pair<EIG_THREADS>(S, sw, n, tphase ? acc_t : nullptr);
    unsigned long long t3 = tphase ? wall_clock64() : 0;
    if (tphase) acc_t[2] += t3 - t2;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (!sw[q]) continue;
      for (int k = l[q] + tid; k < ihi[q]; k += EIG_THREADS) {
        log_c[q][nlog[q] * n + k] = S[q].crot[k];
        log_s[q][nlog[q] * n + k] = S[q].srot[k];
      }
      if (tid == 0) {
        s_lo[q][nlog[q]] = l[q];
        s_hi[q][nlog[q]] = ihi[q];
      }
    }
    __syncthreads();
    unsigned long long t4 = tphase ? wall_clock64() : 0;
    if (tphase) acc_t[3] += t4 - t3;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (sw[q] && ++nlog[q] == EIG_KLOG) {
        flush_log<EIG_THREADS>(S[q].H, Zt[q], n, nlog[q], s_lo[q], s_hi[q], log_c[q], log_s[q], S[q].hb);
        nlog[q] = 0;
      }
    }
    if (tphase) acc_t[4] += wall_clock64() - t4;
  }
}

#define KHX_EIG_BOUNDS(T) __launch_bounds__(T, T >= 256 ? 1024 / T : 1)
#define KHX_EIG_ARGS                                                                                   \
  int n, cd *A, long long sA, cd *Ztg, cd *Xg, long long sW, cd *lam, long long sL, int *info,        \
      unsigned long long *tphase, int flags
// n <= KHX_EIG_LDS_N: the whole decomposition in LDS
template <int EIG_THREADS>
__global__ KHX_EIG_BOUNDS(EIG_THREADS) void zeig_schur_kernel(KHX_EIG_ARGS) {
  eig_body<EIG_THREADS, 7>(n, A, sA, Ztg, Xg, sW, lam, sL, info, tphase, flags);
}
// n > KHX_EIG_LDS_N: one stage per launch
template <int EIG_THREADS>
__global__ KHX_EIG_BOUNDS(EIG_THREADS) void zeig_hess_kernel(KHX_EIG_ARGS) {
  eig_body<EIG_THREADS, 1>(n, A, sA, Ztg, Xg, sW, lam, sL, info, tphase, flags);
}
// The Schur stage: single-shift sweeps (CHAIN = false, four 256-thread
// workgroups per CU), or chain passes of MS_NS shifts on large active blocks
// (CHAIN = true: they need more registers than four workgroups per CU
// leave, QRW = workgroups per CU the kernel is compiled for).
template <int EIG_THREADS, int QRW, bool CHAIN>
__global__ __launch_bounds__(EIG_THREADS, QRW) void zeig_qr_kernel(KHX_EIG_ARGS) {
  eig_body<EIG_THREADS, 2, CHAIN>(n, A, sA, Ztg, Xg, sW, lam, sL, info, tphase, flags);
}
// the default launches (single-shift sweeps): four waves per SIMD, i.e. at
// most 128 registers (256 threads: four workgroups per CU; 128 threads:
// eight, when the LDS allows)
__global__ __launch_bounds__(256, 4) __attribute__((amdgpu_num_vgpr(128))) void zeig_qr256_kernel(KHX_EIG_ARGS) {
  eig_body<256, 2, false>(n, A, sA, Ztg, Xg, sW, lam, sL, info, tphase, flags);
}
__global__ __launch_bounds__(128, 4) __attribute__((amdgpu_num_vgpr(128))) void zeig_qr128_kernel(KHX_EIG_ARGS) {
  eig_body<128, 2, false>(n, A, sA, Ztg, Xg, sW, lam, sL, info, tphase, flags);
}
// two matrices per workgroup (eig_qr_pair): four waves per SIMD, i.e. at
// most 128 registers
template <int EIG_THREADS>
__global__ __launch_bounds__(EIG_THREADS, 4) __attribute__((amdgpu_num_vgpr(128))) void zeig_qr_pair_kernel(
    KHX_EIG_ARGS, int batch) {
  eig_qr_pair<EIG_THREADS>(n, A, sA, Ztg, Xg, sW, lam, sL, info, batch, tphase, flags);
}
template <int EIG_THREADS>
__global__ KHX_EIG_BOUNDS(EIG_THREADS) void zeig_vec_kernel(KHX_EIG_ARGS) {
  eig_body<EIG_THREADS, 4>(n, A, sA, Ztg, Xg, sW, lam, sL, info, tphase, flags);
}

// Column normalisation of W: unit 2-norm, largest |component| made real.
// rscale (optional, n per item): 1 / the factor applied to each column.
__global__ __launch_bounds__(256) void znormalize_cols_kernel(int n, cd* W, long long sW, cd* rscale) {
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
    if (rscale) rscale[item * n + k] = sqrt(ss) * conjc(rot);  // 1 / f (|rot| = 1)
  }
}

size_t zeig_workspace_bytes(int n, int batch) {
  return 2ull * (size_t)n * n * sizeof(cd) * (size_t)batch;  // Zt and X
}

static size_t eig_lds_bytes(int n, int threads, bool chain = false) {
  return (size_t)(eig_r1_elems(n, chain) + eig_wv_elems(n) + eig_part_elems(n, threads / 64, chain) +
                  eig_tail_elems(n, chain)) *
         sizeof(cd);
}
static bool eig_chain_on() {
  const char* ms = getenv("KHX_EIG_MULTISHIFT");
  return ms && atoi(ms) != 0;
}

// The Schur-stage kernel for T threads: single-shift sweeps by default;
// KHX_EIG_MULTISHIFT=1 selects the chain passes (256 threads: at
// KHX_EIG_QR_WG workgroups per CU, default 2).
template <int T>
static const void* eig_qr_fn() {
  constexpr int W = T >= 256 ? 1024 / T : 1;
  const char* ms = getenv("KHX_EIG_MULTISHIFT");
  if (!(ms && atoi(ms) != 0))
    return T == 256   ? (const void*)zeig_qr256_kernel
           : T == 128 ? (const void*)zeig_qr128_kernel
                      : (const void*)zeig_qr_kernel<T, W, false>;
  if (T != 256) return (const void*)zeig_qr_kernel<T, (T > 256 ? 1024 / T : 1), true>;
  const char* env = getenv("KHX_EIG_QR_WG");
  const int w = env ? atoi(env) : 2;
  return w == 4 ? (const void*)zeig_qr_kernel<256, 4, true>
         : w == 3 ? (const void*)zeig_qr_kernel<256, 3, true>
                  : (const void*)zeig_qr_kernel<256, 2, true>;
}

// Paired Schur stage (zeig_qr_pair_kernel) for the default single-shift
// sweeps of a 256-thread launch, while two workgroups fit a CU's LDS:
// KHX_EIG_PAIR=1 (256 threads per pair) or 2 (512).  Returns the pair
// kernel's thread count, 0 for the one-matrix kernel.
static size_t eig_pair_lds_bytes(int n) { return 2 * (size_t)eig_pair_elems(n) * sizeof(cd); }
static int eig_pair_threads(int n, int T, int tq, int flags, size_t pad) {
  const char* env = getenv("KHX_EIG_PAIR");
  const int v = env ? atoi(env) : 0;
  if (v != 1 && v != 2) return 0;
  const bool ok = n > KHX_EIG_LDS_N && eig_pair_lds_bytes(n) <= 80 * 1024 && T == 256 && tq == 256 &&
                  !eig_chain_on() && (flags & 4) && pad == 0;
  return ok ? (v == 2 ? 512 : 256) : 0;
}
static const void* eig_pair_fn(int threads) {
  return threads == 512 ? (const void*)zeig_qr_pair_kernel<512> : (const void*)zeig_qr_pair_kernel<256>;
}

// Threads per matrix in the Schur stage: those of the other stages, or 128
// with KHX_EIG_QR_THREADS=128 (single-shift sweeps only).
static int eig_qr_threads(int T) {
  const char* env = getenv("KHX_EIG_QR_THREADS");
  const char* ms = getenv("KHX_EIG_MULTISHIFT");
  if (env && atoi(env) == 128 && !(ms && atoi(ms) != 0)) return 128;
  return T;
}

// One workgroup per matrix: n <= KHX_EIG_LDS_N in one launch, else the
// Hessenberg, Schur and eigenvector stages in three launches.
template <int T>
static int launch_eig(hipStream_t s, int n, int batch, size_t lds, Mat A, cd* Zt, cd* X, long long sZ, cd* lam,
                      long long lstride, int* info, unsigned long long* tphase, int flags) {
  if (n <= KHX_EIG_LDS_N) {
    KHX_TRY(ensure_dynamic_lds((const void*)zeig_schur_kernel<T>, lds));
    zeig_schur_kernel<T><<<dim3(batch), dim3(T), lds, s>>>(n, A.p, A.stride, Zt, X, sZ, lam, lstride, info, tphase,
                                                           flags);
    KHX_LAUNCH_CHECK();
    return 0;
  }
  const int tq = eig_qr_threads(T);
  const void* qr = tq == 128 ? eig_qr_fn<128>() : eig_qr_fn<T>();
  const size_t pad = lds - eig_lds_bytes(n, T);  // KHX_EIG_LDS_PAD
  const size_t ldsq = eig_lds_bytes(n, tq, eig_chain_on()) + pad;
  KHX_TRY(ensure_dynamic_lds((const void*)zeig_hess_kernel<T>, lds));
  KHX_TRY(ensure_dynamic_lds(qr, ldsq));
  KHX_TRY(ensure_dynamic_lds((const void*)zeig_vec_kernel<T>, lds));
  // Hessenberg panel in LDS (flags bit 3) where it fits the launch's LDS, or
  // where a larger LDS costs nothing (512 or more threads: one matrix per CU)
  const int hw = T / 64;
  const size_t ldsh_need =
      lds + (size_t)std::max(0, hess_part_elems(n, hw) - eig_part_elems(n, hw, false)) * sizeof(cd);
  const char* hl_env = getenv("KHX_EIG_HESS_LDS");
  const bool hess_lds = !(hl_env && atoi(hl_env) == 0) && (ldsh_need == lds || T >= 512) && ldsh_need <= 150 * 1024;
  const size_t ldsh = hess_lds ? ldsh_need : lds;
  const int hflags = flags | (hess_lds ? 8 : 0);
  if (ldsh != lds) KHX_TRY(ensure_dynamic_lds((const void*)zeig_hess_kernel<T>, ldsh));
  const int pt = eig_pair_threads(n, T, tq, flags, pad);
  if (pt) KHX_TRY(ensure_dynamic_lds(eig_pair_fn(pt), eig_pair_lds_bytes(n)));
  zeig_hess_kernel<T><<<dim3(batch), dim3(T), ldsh, s>>>(n, A.p, A.stride, Zt, X, sZ, lam, lstride, info, tphase,
                                                         hflags);
  KHX_LAUNCH_CHECK();
  if (pt) {
    size_t ldsp = eig_pair_lds_bytes(n);
    void* args[] = {&n, &A.p, &A.stride, &Zt, &X, &sZ, &lam, &lstride, &info, &tphase, &flags, &batch};
    KHX_CHECK_HIP(hipLaunchKernel(eig_pair_fn(pt), dim3((batch + 1) / 2), dim3(pt), args, ldsp, s));
  } else {
    void* args[] = {&n, &A.p, &A.stride, &Zt, &X, &sZ, &lam, &lstride, &info, &tphase, &flags};
    KHX_CHECK_HIP(hipLaunchKernel(qr, dim3(batch), dim3(tq), args, ldsq, s));
  }
  zeig_vec_kernel<T><<<dim3(batch), dim3(T), lds, s>>>(n, A.p, A.stride, Zt, X, sZ, lam, lstride, info, tphase, flags);
  KHX_LAUNCH_CHECK();
  return 0;
}

// Matrices the eigensolver keeps in flight at once on the current device
// (resident workgroups per CU x CUs) with the default launch shape: a batch
// that is a multiple of this fills every round of workgroups.
int zeig_slots(int n) {
  KHX_REQUIRE(n >= 1 && n <= 8192, "zeig_slots: bad order");
  const char* env = getenv("KHX_EIG_THREADS");
  const int threads = env ? atoi(env) : EIG_THREADS_DEFAULT;
  const char* pad_env = getenv("KHX_EIG_LDS_PAD");
  const size_t lds = eig_lds_bytes(n, threads, n > KHX_EIG_LDS_N && eig_chain_on()) + (pad_env ? (size_t)atol(pad_env) : 0);
  KHX_REQUIRE(lds <= 150 * 1024, "zeig_slots: n too large for the LDS work vectors");
  // the Schur stage (the longest) decides how many matrices are in flight
  const bool small = n <= KHX_EIG_LDS_N;
  if (!small && eig_qr_threads(threads) == 128) {
    const void* f = eig_qr_fn<128>();
    const size_t l = eig_lds_bytes(n, 128, false);
    KHX_TRY(ensure_dynamic_lds(f, l));
    int per_cu = 0, dev = 0, cus = 0;
    KHX_CHECK_HIP(hipGetDevice(&dev));
    KHX_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    KHX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 128, l));
    return per_cu * cus > 0 ? per_cu * cus : 1;
  }
  {
    const char* rc_env = getenv("KHX_EIG_REG_CHASE");
    const int flags = (rc_env && atoi(rc_env) == 0) ? 0 : 4;
    const size_t pad = pad_env ? (size_t)atol(pad_env) : 0;
    const int pt = small ? 0 : eig_pair_threads(n, threads, eig_qr_threads(threads), flags, pad);
    if (pt) {
      const void* f = eig_pair_fn(pt);
      const size_t l = eig_pair_lds_bytes(n);
      KHX_TRY(ensure_dynamic_lds(f, l));
      int per_cu = 0, dev = 0, cus = 0;
      KHX_CHECK_HIP(hipGetDevice(&dev));
      KHX_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      KHX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, pt, l));
      return 2 * per_cu * cus > 0 ? 2 * per_cu * cus : 1;
    }
  }
  const void* fn = threads == 1024 ? (small ? (const void*)zeig_schur_kernel<1024> : eig_qr_fn<1024>())
                   : threads == 512 ? (small ? (const void*)zeig_schur_kernel<512> : eig_qr_fn<512>())
                   : threads == 128 ? (small ? (const void*)zeig_schur_kernel<128> : eig_qr_fn<128>())
                   : threads == 64  ? (small ? (const void*)zeig_schur_kernel<64> : eig_qr_fn<64>())
                                    : (small ? (const void*)zeig_schur_kernel<256> : eig_qr_fn<256>());
  KHX_TRY(ensure_dynamic_lds(fn, lds));
  int per_cu = 0, dev = 0, cus = 0;
  KHX_CHECK_HIP(hipGetDevice(&dev));
  KHX_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  KHX_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds));
  return per_cu * cus > 0 ? per_cu * cus : 1;
}

int zeig(hipStream_t s, int n, Mat A, cd* lam, long long lstride, Mat W, int batch, int* info, void* ws,
         unsigned long long* tphase, cd* rscale) {
  if (batch <= 0) return 0;
  KHX_REQUIRE(n >= 1 && n <= 8192, "zeig: bad order");
  KHX_REQUIRE(A.ld == n && W.ld == n && A.stride != 0 && W.stride != 0,
              "zeig: matrices must be dense (ld == n) and not broadcast");
  cd* Zt = static_cast<cd*>(ws);
  cd* X = Zt + (size_t)n * n * batch;
  const long long sZ = (long long)n * n;
  const char* env = getenv("KHX_EIG_THREADS");
  // a batch that leaves CUs idle gets 512 threads per matrix (measured faster
  // there: the blocked Hessenberg and the deferred passes spread over more lanes)
  int threads = EIG_THREADS_DEFAULT;
  if (env) {
    threads = atoi(env);
  } else {
    static int cus[16] = {0};
    int dev = 0;
    KHX_CHECK_HIP(hipGetDevice(&dev));
    if (dev >= 0 && dev < 16) {
      if (cus[dev] == 0) KHX_CHECK_HIP(hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev));
      if (batch <= cus[dev]) threads = 512;
    }
  }
  KHX_REQUIRE(threads == 64 || threads == 128 || threads == 256 || threads == 512 || threads == 1024,
              "zeig: KHX_EIG_THREADS must be 64, 128, 256, 512 or 1024");
  // optional LDS padding: fewer resident workgroups per CU keep the working
  // set of the resident matrices inside the caches
  const char* pad_env = getenv("KHX_EIG_LDS_PAD");
  // bit 0: chase each LDS block on one wavefront (default on; KHX_EIG_CHASE_WAVE=0 turns it off)
  const char* cw_env = getenv("KHX_EIG_CHASE_WAVE");
  // bit 2: register chase (qr_sweep_reg; default on, KHX_EIG_REG_CHASE=0 selects the LDS chase)
  const char* rc_env = getenv("KHX_EIG_REG_CHASE");
  // bit 1: row-by-row eigenvector sweep (KHX_EIG_VEC_ROWS=1) instead of one thread per eigenvector
  const char* vr_env = getenv("KHX_EIG_VEC_ROWS");
  const char* cap_env = getenv("KHX_EIG_ITMAX");
  const int cap = cap_env ? atoi(cap_env) : 0;
  KHX_REQUIRE(cap >= 0 && cap < (1 << 22), "zeig: KHX_EIG_ITMAX out of range");
  const int flags = ((cw_env && atoi(cw_env) == 0) ? 0 : 1) | ((vr_env && atoi(vr_env) != 0) ? 2 : 0) |
                    ((rc_env && atoi(rc_env) == 0) ? 0 : 4) | (cap << 8);
  const size_t lds = eig_lds_bytes(n, threads) + (pad_env ? (size_t)atol(pad_env) : 0);
  KHX_REQUIRE(lds <= 150 * 1024, "zeig: n too large for the LDS work vectors");
  {
    ProfScope ps("zeig", s);
    switch (threads) {
      case 1024: KHX_TRY(launch_eig<1024>(s, n, batch, lds, A, Zt, X, sZ, lam, lstride, info, tphase, flags)); break;
      case 512: KHX_TRY(launch_eig<512>(s, n, batch, lds, A, Zt, X, sZ, lam, lstride, info, tphase, flags)); break;
      case 128: KHX_TRY(launch_eig<128>(s, n, batch, lds, A, Zt, X, sZ, lam, lstride, info, tphase, flags)); break;
      case 64: KHX_TRY(launch_eig<64>(s, n, batch, lds, A, Zt, X, sZ, lam, lstride, info, tphase, flags)); break;
      default: KHX_TRY(launch_eig<256>(s, n, batch, lds, A, Zt, X, sZ, lam, lstride, info, tphase, flags)); break;
    }
  }
  // W = Z X = Zt^T X
  KHX_TRY(zgemm_ta(s, n, n, n, mkc(1.0, 0.0), Mat{Zt, n, sZ}, Mat{X, n, sZ}, mkc(0.0, 0.0), W, batch));
  znormalize_cols_kernel<<<dim3(batch), dim3(256), 0, s>>>(n, W.p, W.stride, rscale);
  KHX_LAUNCH_CHECK();
  return 0;
}

// ----------------------------------------------------------------- no-alloc
// khx_eig for callers that cannot allocate: the workspace of zeig (Zt and X,
// 2 n^2 complex per matrix) lives in this file-scope device array, one copy
// per GPU, held until the process ends.  Host code never passes the array's
// name to a kernel: its device address comes from hipGetSymbolAddress, looked
// up once per device with that device current.
constexpr size_t EIG_NOALLOC_BYTES = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[EIG_NOALLOC_BYTES];

size_t zeig_noalloc_limit() { return EIG_NOALLOC_BYTES; }

static int scratch_address(void** out) {
  static std::mutex mu;
  static std::vector<void*> ptr;  // per device, nullptr until looked up
  int dev = 0, ndev = 0;
  KHX_CHECK_HIP(hipGetDevice(&dev));
  KHX_CHECK_HIP(hipGetDeviceCount(&ndev));
  std::lock_guard<std::mutex> lock(mu);
  if ((int)ptr.size() < ndev) ptr.resize(ndev, nullptr);
  KHX_REQUIRE(dev >= 0 && dev < ndev, "zeig_noalloc: bad current device");
  if (!ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return -3;
    }
    ptr[dev] = p;
  }
  *out = ptr[dev];
  return 0;
}

int zeig_noalloc(hipStream_t s, int n, Mat A, cd* lam, long long lstride, Mat W, int batch, int* info) {
  if (batch <= 0) return 0;
  KHX_REQUIRE(n >= 1 && n <= 8192, "khx_eig_noalloc: bad order");
  KHX_REQUIRE(zeig_workspace_bytes(n, batch) <= EIG_NOALLOC_BYTES,
              "khx_eig_noalloc: batch x order exceeds the static workspace (2 n^2 x 16 bytes per matrix, at most "
              "64 MiB)");
  void* ws = nullptr;
  KHX_TRY(scratch_address(&ws));
  return zeig(s, n, A, lam, lstride, W, batch, info, ws);
}

}  // namespace khx
