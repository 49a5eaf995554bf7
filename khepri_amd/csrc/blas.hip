// Batched complex FP64 building blocks: GEMM, LU solve, elementwise helpers.
//
// All matrices are row-major complex128 (numpy layout), batch-major; a batch
// stride of 0 broadcasts one operand to every item (e.g. the convolution
// matrix of a layer, shared by all frequencies and Bloch vectors).
#include "khx_internal.h"
#include "khx_prof.h"

#include <array>
#include <cstdlib>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace khx {

// ======================================================================== GEMM
// One workgroup = one BM x BN tile of one batch item; 256 threads arranged
// (BM/TM) x (BN/TN), each owning a TM x TN register tile.  Operand tiles are
// staged in LDS, A transposed so a wave reads a few broadcast rows of A and
// consecutive columns of B.  OA / OB select op(X) = X (0), X^T (1) or X^H (2);
// a transposed operand is stored with the other dimension first.
template <int BM, int BN, int BK, int TM, int TN, int OA, int OB>
__global__ __launch_bounds__(256) void zgemm_kernel(int M, int N, int K, cd alpha,
                                                    const cd* __restrict__ A, int lda, long long sA,
                                                    const cd* __restrict__ B, int ldb, long long sB,
                                                    cd beta, cd* __restrict__ C, int ldc, long long sC,
                                                    int tilesN, int tilesPerItem) {
  constexpr int TX = BN / TN;  // threads along N
  constexpr int TY = BM / TM;  // threads along M
  static_assert(TX * TY == 256, "256 threads");
  __shared__ cd As[BK][BM];
  __shared__ cd Bs[BK][BN];

  const long long item = blockIdx.x / tilesPerItem;
  const int tile = blockIdx.x % tilesPerItem;
  const int row0 = (tile / tilesN) * BM;
  const int col0 = (tile % tilesN) * BN;
  const cd* a = A + item * sA;
  const cd* b = B + item * sB;
  cd* c = C + item * sC;

  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;

  cd acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = mkc(0.0, 0.0);

  for (int k0 = 0; k0 < K; k0 += BK) {
    // stage A tile (BM x BK) -> As[k][m]; op(A) = A^T reads A's rows along m
#pragma unroll
    for (int e = tid; e < BM * BK; e += 256) {
      const int r = OA ? e % BM : e / BK, kk = OA ? e / BM : e % BK;
      const int gr = row0 + r, gk = k0 + kk;
      const long long off = OA ? (long long)gk * lda + gr : (long long)gr * lda + gk;
      cd v = (gr < M && gk < K) ? a[off] : mkc(0.0, 0.0);
      if (OA == 2) v.y = -v.y;
      As[kk][r] = v;
    }
    // stage B tile (BK x BN)
#pragma unroll
    for (int e = tid; e < BK * BN; e += 256) {
      const int kk = OB ? e % BK : e / BN, cc = OB ? e / BK : e % BN;
      const int gk = k0 + kk, gc = col0 + cc;
      const long long off = OB ? (long long)gc * ldb + gk : (long long)gk * ldb + gc;
      cd v = (gk < K && gc < N) ? b[off] : mkc(0.0, 0.0);
      if (OB == 2) v.y = -v.y;
      Bs[kk][cc] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < BK; ++kk) {
      cd av[TM], bv[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) av[i] = As[kk][ty + TY * i];
#pragma unroll
      for (int j = 0; j < TN; ++j) bv[j] = Bs[kk][tx + TX * j];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = fmac(acc[i][j], av[i], bv[j]);
    }
    __syncthreads();
  }

  const bool use_beta = (beta.x != 0.0 || beta.y != 0.0);
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int gr = row0 + ty + TY * i;
    if (gr >= M) continue;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int gc = col0 + tx + TX * j;
      if (gc >= N) continue;
      cd v = alpha * acc[i][j];
      cd* dst = c + (long long)gr * ldc + gc;
      if (use_beta) v = fmac(v, beta, *dst);
      *dst = v;
    }
  }
}

template <int BM, int BN, int BK, int TM, int TN, int OA, int OB>
static int launch_zgemm(hipStream_t s, int M, int N, int K, cd alpha, Mat A, Mat B, cd beta, Mat C,
                        int batch) {
  const int tilesM = (M + BM - 1) / BM, tilesN = (N + BN - 1) / BN;
  const long long blocks = (long long)tilesM * tilesN * batch;
  KHX_REQUIRE(blocks < (1ll << 31), "zgemm: grid too large");
  ProfScope ps("zgemm_kernel", s);
  zgemm_kernel<BM, BN, BK, TM, TN, OA, OB><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(
      M, N, K, alpha, A.p, A.ld, A.stride, B.p, B.ld, B.stride, beta, C.p, C.ld, C.stride, tilesN,
      tilesM * tilesN);
  KHX_LAUNCH_CHECK();
  return 0;
}

// ------------------------------------------------------------- MFMA GEMM
// The same product on the FP64 matrix cores: v_mfma_f64_16x16x4f64 computes a
// 16 x 16 block D += A (16 x 4) B (4 x 16); lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16] and holds D[l / 16 + 4 r][l % 16] in register r.
// A complex product is four real ones: Re += Ar Br - Ai Bi, Im += Ar Bi + Ai Br.
// A workgroup (4 wavefronts in a 2 x 2 grid) owns a (32 WM) x (32 WN) tile of
// C; each wavefront owns WM x WN blocks of 16 x 16.  K advances 16 at a time
// through LDS (k-major, one complex of padding per row against bank
// conflicts); the next K slice is fetched into registers while the current
// one is multiplied.  Consecutive tiles of one batch item are kept on one XCD
// (workgroups are dealt round-robin to the 8 XCDs) so they share its L2.
typedef double khx_d4 __attribute__((ext_vector_type(4)));

template <int WM, int WN, int OA, int OB>
__global__ __launch_bounds__(256) void zgemm_mfma_kernel(int M, int N, int K, cd alpha,
                                                         const cd* __restrict__ A, int lda, long long sA,
                                                         const cd* __restrict__ B, int ldb, long long sB,
                                                         cd beta, cd* __restrict__ C, int ldc, long long sC,
                                                         int tilesN, int tilesPerItem, int xcd_remap) {
  constexpr int BM = 32 * WM, BN = 32 * WN, BK = 16;
  constexpr int LA = (BM * BK) / 256, LBn = (BK * BN) / 256;  // complex loads per thread
  __shared__ cd As[BK][BM + 1];
  __shared__ cd Bs[BK][BN + 1];

  unsigned bid = blockIdx.x;
  if (xcd_remap) {  // gridDim.x % 8 == 0: logical tile order = XCD-major
    const unsigned per = gridDim.x / 8;
    bid = (bid % 8) * per + bid / 8;
  }
  const long long item = bid / tilesPerItem;
  const int tile = bid % tilesPerItem;
  const int row0 = (tile / tilesN) * BM;
  const int col0 = (tile % tilesN) * BN;
  const cd* a = A + item * sA;
  const cd* b = B + item * sB;
  cd* c = C + item * sC;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * (16 * WM), wc = (wave & 1) * (16 * WN);
  const int li = lane & 15, lk = lane >> 4;

  khx_d4 accR[WM][WN], accI[WM][WN];
#pragma unroll
  for (int t = 0; t < WM; ++t)
#pragma unroll
    for (int u = 0; u < WN; ++u) {
      accR[t][u] = khx_d4{0.0, 0.0, 0.0, 0.0};
      accI[t][u] = khx_d4{0.0, 0.0, 0.0, 0.0};
    }

  cd ra[LA], rb[LBn];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const int e = tid + 256 * q;
      const int r = OA ? e % BM : e / BK, kk = OA ? e / BM : e % BK;
      const int gr = row0 + r, gk = k0 + kk;
      const long long off = OA ? (long long)gk * lda + gr : (long long)gr * lda + gk;
      cd v = (gr < M && gk < K) ? a[off] : mkc(0.0, 0.0);
      if (OA == 2) v.y = -v.y;
      ra[q] = v;
    }
#pragma unroll
    for (int q = 0; q < LBn; ++q) {
      const int e = tid + 256 * q;
      const int kk = OB ? e % BK : e / BN, cc = OB ? e / BK : e % BN;
      const int gk = k0 + kk, gc = col0 + cc;
      const long long off = OB ? (long long)gc * ldb + gk : (long long)gk * ldb + gc;
      cd v = (gk < K && gc < N) ? b[off] : mkc(0.0, 0.0);
      if (OB == 2) v.y = -v.y;
      rb[q] = v;
    }
  };

  fetch(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const int e = tid + 256 * q;
      const int r = OA ? e % BM : e / BK, kk = OA ? e / BM : e % BK;
      As[kk][r] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < LBn; ++q) {
      const int e = tid + 256 * q;
      const int kk = OB ? e % BK : e / BN, cc = OB ? e / BK : e % BN;
      Bs[kk][cc] = rb[q];
    }
    __syncthreads();
    if (k0 + BK < K) fetch(k0 + BK);
#pragma unroll
    for (int kq = 0; kq < BK / 4; ++kq) {
      cd av[WM], bv[WN];
#pragma unroll
      for (int t = 0; t < WM; ++t) av[t] = As[kq * 4 + lk][wr + t * 16 + li];
#pragma unroll
      for (int u = 0; u < WN; ++u) bv[u] = Bs[kq * 4 + lk][wc + u * 16 + li];
#pragma unroll
      for (int t = 0; t < WM; ++t) {
        const double nai = -av[t].y;
#pragma unroll
        for (int u = 0; u < WN; ++u) {
          accR[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t].x, bv[u].x, accR[t][u], 0, 0, 0);
          accR[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, bv[u].y, accR[t][u], 0, 0, 0);
          accI[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t].x, bv[u].y, accI[t][u], 0, 0, 0);
          accI[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t].y, bv[u].x, accI[t][u], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }

  const bool use_beta = (beta.x != 0.0 || beta.y != 0.0);
#pragma unroll
  for (int t = 0; t < WM; ++t)
#pragma unroll
    for (int u = 0; u < WN; ++u) {
      const int gc = col0 + wc + u * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = row0 + wr + t * 16 + lk + 4 * r;
        if (gr >= M || gc >= N) continue;
        cd v = alpha * mkc(accR[t][u][r], accI[t][u][r]);
        cd* dst = c + (long long)gr * ldc + gc;
        if (use_beta) v = fmac(v, beta, *dst);
        *dst = v;
      }
    }
}

template <int WM, int WN, int OA, int OB>
static int launch_zgemm_mfma(hipStream_t s, int M, int N, int K, cd alpha, Mat A, Mat B, cd beta, Mat C,
                             int batch) {
  constexpr int BM = 32 * WM, BN = 32 * WN;
  const int tilesM = (M + BM - 1) / BM, tilesN = (N + BN - 1) / BN;
  const long long blocks = (long long)tilesM * tilesN * batch;
  KHX_REQUIRE(blocks < (1ll << 31), "zgemm: grid too large");
  ProfScope ps("zgemm_kernel", s);
  zgemm_mfma_kernel<WM, WN, OA, OB><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(
      M, N, K, alpha, A.p, A.ld, A.stride, B.p, B.ld, B.stride, beta, C.p, C.ld, C.stride, tilesN,
      tilesM * tilesN, (blocks % 8 == 0) ? 1 : 0);
  KHX_LAUNCH_CHECK();
  return 0;
}

// ------------------------------------------------------------- FP32 MFMA GEMM (mixed precision)
// C = alpha A B + beta C (no transposes) with the product on the FP32 matrix
// cores (v_mfma_f32_16x16x4f32, the same tiling as zgemm_mfma_kernel): the
// complex-double operands are rounded to complex float when they are staged
// in LDS, partial sums are FP32 and the epilogue (alpha, beta, C) is FP64.
// The product's relative error is about 2^-24 (K = 16..324 here), so it is
// used only where an FP64 residual refinement follows: the triangular-solve
// updates of the mixed-precision zgesv (KHX_LU_MIXED=1).
typedef float khx_f4 __attribute__((ext_vector_type(4)));
struct cfl {
  float x, y;
};
// v_mfma_f32_16x16x4f32 output: lane l holds D[F32_ROW(l / 16, r)][l % 16] in register r
#define F32_ROW(lk, r) (4 * (lk) + (r))

template <int WM, int WN>
__global__ __launch_bounds__(256) void zgemm_f32mfma_kernel(int M, int N, int K, cd alpha,
                                                            const cd* __restrict__ A, int lda, long long sA,
                                                            const cd* __restrict__ B, int ldb, long long sB,
                                                            cd beta, cd* __restrict__ C, int ldc, long long sC,
                                                            int tilesN, int tilesPerItem, int xcd_remap) {
  constexpr int BM = 32 * WM, BN = 32 * WN, BK = 16;
  constexpr int LA = (BM * BK) / 256, LBn = (BK * BN) / 256;
  __shared__ cfl As[BK][BM + 1];
  __shared__ cfl Bs[BK][BN + 1];
  unsigned bid = blockIdx.x;
  if (xcd_remap) {
    const unsigned per = gridDim.x / 8;
    bid = (bid % 8) * per + bid / 8;
  }
  const long long item = bid / tilesPerItem;
  const int tile = bid % tilesPerItem;
  const int row0 = (tile / tilesN) * BM;
  const int col0 = (tile % tilesN) * BN;
  const cd* a = A + item * sA;
  const cd* b = B + item * sB;
  cd* c = C + item * sC;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * (16 * WM), wc = (wave & 1) * (16 * WN);
  const int li = lane & 15, lk = lane >> 4;
  khx_f4 accR[WM][WN], accI[WM][WN];
#pragma unroll
  for (int t = 0; t < WM; ++t)
#pragma unroll
    for (int u = 0; u < WN; ++u) {
      accR[t][u] = khx_f4{0.f, 0.f, 0.f, 0.f};
      accI[t][u] = khx_f4{0.f, 0.f, 0.f, 0.f};
    }
  cfl ra[LA], rb[LBn];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const int e = tid + 256 * q;
      const int r = e / BK, kk = e % BK;
      const int gr = row0 + r, gk = k0 + kk;
      const cd v = (gr < M && gk < K) ? a[(long long)gr * lda + gk] : mkc(0.0, 0.0);
      ra[q] = cfl{(float)v.x, (float)v.y};
    }
#pragma unroll
    for (int q = 0; q < LBn; ++q) {
      const int e = tid + 256 * q;
      const int kk = e / BN, cc = e % BN;
      const int gk = k0 + kk, gc = col0 + cc;
      const cd v = (gk < K && gc < N) ? b[(long long)gk * ldb + gc] : mkc(0.0, 0.0);
      rb[q] = cfl{(float)v.x, (float)v.y};
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const int e = tid + 256 * q;
      As[e % BK][e / BK] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < LBn; ++q) {
      const int e = tid + 256 * q;
      Bs[e / BN][e % BN] = rb[q];
    }
    __syncthreads();
    if (k0 + BK < K) fetch(k0 + BK);
#pragma unroll
    for (int kq = 0; kq < BK / 4; ++kq) {
      cfl av[WM], bv[WN];
#pragma unroll
      for (int t = 0; t < WM; ++t) av[t] = As[kq * 4 + lk][wr + t * 16 + li];
#pragma unroll
      for (int u = 0; u < WN; ++u) bv[u] = Bs[kq * 4 + lk][wc + u * 16 + li];
#pragma unroll
      for (int t = 0; t < WM; ++t) {
        const float nai = -av[t].y;
#pragma unroll
        for (int u = 0; u < WN; ++u) {
          accR[t][u] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[t].x, bv[u].x, accR[t][u], 0, 0, 0);
          accR[t][u] = __builtin_amdgcn_mfma_f32_16x16x4f32(nai, bv[u].y, accR[t][u], 0, 0, 0);
          accI[t][u] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[t].x, bv[u].y, accI[t][u], 0, 0, 0);
          accI[t][u] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[t].y, bv[u].x, accI[t][u], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  const bool use_beta = (beta.x != 0.0 || beta.y != 0.0);
#pragma unroll
  for (int t = 0; t < WM; ++t)
#pragma unroll
    for (int u = 0; u < WN; ++u) {
      const int gc = col0 + wc + u * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = row0 + wr + t * 16 + F32_ROW(lk, r);
        if (gr >= M || gc >= N) continue;
        cd v = alpha * mkc((double)accR[t][u][r], (double)accI[t][u][r]);
        cd* dst = c + (long long)gr * ldc + gc;
        if (use_beta) v = fmac(v, beta, *dst);
        *dst = v;
      }
    }
}

int zgemm_f32(hipStream_t s, int M, int N, int K, cd alpha, Mat A, Mat B, cd beta, Mat C, int batch) {
  if (batch <= 0 || M <= 0 || N <= 0) return 0;
  const bool small = M <= 96 || N <= 96;
  const int BM = 32, BN = small ? 32 : 64;
  const int tilesM = (M + BM - 1) / BM, tilesN = (N + BN - 1) / BN;
  const long long blocks = (long long)tilesM * tilesN * batch;
  KHX_REQUIRE(blocks < (1ll << 31), "zgemm_f32: grid too large");
  ProfScope ps("zgemm_kernel", s);
  if (small)
    zgemm_f32mfma_kernel<1, 1><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(
        M, N, K, alpha, A.p, A.ld, A.stride, B.p, B.ld, B.stride, beta, C.p, C.ld, C.stride, tilesN,
        tilesM * tilesN, (blocks % 8 == 0) ? 1 : 0);
  else
    zgemm_f32mfma_kernel<1, 2><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(
        M, N, K, alpha, A.p, A.ld, A.stride, B.p, B.ld, B.stride, beta, C.p, C.ld, C.stride, tilesN,
        tilesM * tilesN, (blocks % 8 == 0) ? 1 : 0);
  KHX_LAUNCH_CHECK();
  return 0;
}

// ---- tile selection.  Eight kernels (all 256 threads): FMA tiles
//   0: 64 x 64 (4 x 4 per thread), 1: 32 x 32 (2 x 2), 2: 64 x 32 (4 x 2), 3: 32 x 64 (2 x 4);
// MFMA tiles 4: 64 x 64, 5: 32 x 32, 6: 64 x 32, 7: 32 x 64.
// The default is MFMA 32 x 32 when a side is at most 96 and MFMA 32 x 64 otherwise
// (the fastest of the eight on the batched layer shapes); a per-shape choice installed by the
// autotuner (khepri_amd/autotune.py, through khx_gemm_tune_set) takes precedence, and
// khx_gemm_force overrides both while candidates are timed.  With recording on, every
// distinct (M, N, K, opA, opB, batch class) passing through is kept for the autotuner.
namespace {
std::mutex g_tune_mu;
std::unordered_map<unsigned long long, int> g_tune;
std::vector<std::array<int, 6>> g_shapes;
bool g_record = false;
int g_force = -1;
unsigned long long tune_key(int M, int N, int K, int oa, int ob, int batch) {
  const unsigned long long bc = batch >= 256 ? 2 : (batch >= 16 ? 1 : 0);
  return ((unsigned long long)M << 42) | ((unsigned long long)N << 26) | ((unsigned long long)K << 10) |
         ((unsigned long long)(oa * 3 + ob) << 2) | bc;
}
}  // namespace

int gemm_tile_choice(int M, int N, int K, int oa, int ob, int batch) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (g_force >= 0) return g_force;
  const unsigned long long key = tune_key(M, N, K, oa, ob, batch);
  if (g_record) {
    bool seen = false;
    for (const auto& sh : g_shapes)
      if (tune_key(sh[0], sh[1], sh[2], sh[3], sh[4], sh[5]) == key) seen = true;
    if (!seen && g_shapes.size() < 4096) g_shapes.push_back({M, N, K, oa, ob, batch});
  }
  const auto it = g_tune.find(key);
  if (it != g_tune.end()) return it->second;
  return (M <= 96 || N <= 96) ? 5 : 7;
}

template <int OA, int OB>
static int zgemm_sized(hipStream_t s, int M, int N, int K, cd alpha, Mat A, Mat B, cd beta, Mat C, int batch) {
  if (batch <= 0 || M <= 0 || N <= 0) return 0;
  switch (gemm_tile_choice(M, N, K, OA, OB, batch)) {
    case 1: return launch_zgemm<32, 32, 16, 2, 2, OA, OB>(s, M, N, K, alpha, A, B, beta, C, batch);
    case 2: return launch_zgemm<64, 32, 16, 4, 2, OA, OB>(s, M, N, K, alpha, A, B, beta, C, batch);
    case 3: return launch_zgemm<32, 64, 16, 2, 4, OA, OB>(s, M, N, K, alpha, A, B, beta, C, batch);
    case 4: return launch_zgemm_mfma<2, 2, OA, OB>(s, M, N, K, alpha, A, B, beta, C, batch);
    case 5: return launch_zgemm_mfma<1, 1, OA, OB>(s, M, N, K, alpha, A, B, beta, C, batch);
    case 6: return launch_zgemm_mfma<2, 1, OA, OB>(s, M, N, K, alpha, A, B, beta, C, batch);
    case 7: return launch_zgemm_mfma<1, 2, OA, OB>(s, M, N, K, alpha, A, B, beta, C, batch);
    default: return launch_zgemm<64, 64, 16, 4, 4, OA, OB>(s, M, N, K, alpha, A, B, beta, C, batch);
  }
}

int gemm_tune_set(int M, int N, int K, int oa, int ob, int batch, int cfg) {
  KHX_REQUIRE(cfg >= -1 && cfg <= 7, "gemm tile configuration must be 0..7 (or -1 to clear)");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  const unsigned long long key = tune_key(M, N, K, oa, ob, batch);
  if (cfg < 0)
    g_tune.erase(key);
  else
    g_tune[key] = cfg;
  return 0;
}

int gemm_force(int cfg) {
  KHX_REQUIRE(cfg >= -1 && cfg <= 7, "gemm tile configuration must be 0..7 (or -1)");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_force = cfg;
  return 0;
}

int gemm_record(int on, int* shapes, int max_shapes) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  int count = 0;
  if (shapes)
    for (const auto& sh : g_shapes) {
      if (count >= max_shapes) break;
      for (int q = 0; q < 6; ++q) shapes[count * 6 + q] = sh[q];
      ++count;
    }
  if (on == 1) g_shapes.clear();
  g_record = (on == 1);
  return shapes ? count : (int)g_shapes.size();
}

int zgemm(hipStream_t s, int M, int N, int K, cd alpha, Mat A, Mat B, cd beta, Mat C, int batch) {
  return zgemm_sized<0, 0>(s, M, N, K, alpha, A, B, beta, C, batch);
}

// C = alpha A^T B + beta C (A stored K x M)
int zgemm_ta(hipStream_t s, int M, int N, int K, cd alpha, Mat A, Mat B, cd beta, Mat C, int batch) {
  return zgemm_sized<1, 0>(s, M, N, K, alpha, A, B, beta, C, batch);
}

// C = alpha op(A) op(B) + beta C, op in {0: X, 1: X^T, 2: X^H}
int zgemm_op(hipStream_t s, int opa, int opb, int M, int N, int K, cd alpha, Mat A, Mat B, cd beta, Mat C,
             int batch) {
  KHX_REQUIRE(opa >= 0 && opa <= 2 && opb >= 0 && opb <= 2, "zgemm_op: op must be 0 (N), 1 (T) or 2 (C)");
  switch (opa * 3 + opb) {
    case 0: return zgemm_sized<0, 0>(s, M, N, K, alpha, A, B, beta, C, batch);
    case 1: return zgemm_sized<0, 1>(s, M, N, K, alpha, A, B, beta, C, batch);
    case 2: return zgemm_sized<0, 2>(s, M, N, K, alpha, A, B, beta, C, batch);
    case 3: return zgemm_sized<1, 0>(s, M, N, K, alpha, A, B, beta, C, batch);
    case 4: return zgemm_sized<1, 1>(s, M, N, K, alpha, A, B, beta, C, batch);
    case 5: return zgemm_sized<1, 2>(s, M, N, K, alpha, A, B, beta, C, batch);
    case 6: return zgemm_sized<2, 0>(s, M, N, K, alpha, A, B, beta, C, batch);
    case 7: return zgemm_sized<2, 1>(s, M, N, K, alpha, A, B, beta, C, batch);
    default: return zgemm_sized<2, 2>(s, M, N, K, alpha, A, B, beta, C, batch);
  }
}

// ======================================================================== LU solve
// One workgroup per system.  Gaussian elimination with partial pivoting
// (largest |re|+|im| in the column, first index on ties) applied to A and the
// right-hand sides together, then back substitution on the right-hand sides.
// The pivot row and the column multipliers of each step are staged in LDS.
constexpr int GESV_THREADS = 256;

__global__ __launch_bounds__(GESV_THREADS) void zgesv_kernel(int n, int m, cd* A, int lda, long long sA,
                                                             cd* B, int ldb, long long sB, int* info) {
  extern __shared__ cd smem[];
  cd* prowA = smem;        // n
  cd* prowB = smem + n;    // m
  cd* lcol = prowB + m;    // n
  __shared__ double red_v[GESV_THREADS / 64];
  __shared__ int red_i[GESV_THREADS / 64];
  __shared__ cd red_c[GESV_THREADS / 64];
  __shared__ cd s_akk;

  const long long item = blockIdx.x;
  cd* a = A + item * sA;
  cd* bm = B + item * sB;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  constexpr int NW = GESV_THREADS / 64;
  int first_zero = 0;

  for (int k = 0; k < n; ++k) {
    // ---- pivot search in column k
    double best = -1.0;
    int bi = n;
    cd bv = mkc(0.0, 0.0);
    for (int i = k + tid; i < n; i += GESV_THREADS) {
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
    for (int w = 1; w < NW; ++w) {
      if (red_v[w] > best || (red_v[w] == best && red_i[w] < bi)) {
        best = red_v[w];
        bi = red_i[w];
        bv = red_c[w];
      }
    }
    const cd akk = s_akk;
    int p = bi;
    cd piv = bv;
    if (p >= n) {  // column of NaNs: keep the diagonal
      p = k;
      piv = akk;
    }
    const bool zero = (piv.x == 0.0 && piv.y == 0.0);
    if (zero && first_zero == 0) first_zero = k + 1;
    const cd rp = zero ? mkc(0.0, 0.0) : recipc(piv);

    // ---- swap rows k <-> p; new row k goes to LDS; multipliers for rows > k
    for (int j = k + tid; j < n; j += GESV_THREADS) {
      const cd vk = a[(long long)k * lda + j], vp = a[(long long)p * lda + j];
      if (p != k) {
        a[(long long)k * lda + j] = vp;
        a[(long long)p * lda + j] = vk;
      }
      prowA[j] = vp;
    }
    for (int j = tid; j < m; j += GESV_THREADS) {
      const cd vk = bm[(long long)k * ldb + j], vp = bm[(long long)p * ldb + j];
      if (p != k) {
        bm[(long long)k * ldb + j] = vp;
        bm[(long long)p * ldb + j] = vk;
      }
      prowB[j] = vp;
    }
    for (int i = k + 1 + tid; i < n; i += GESV_THREADS) {
      const cd v = (i == p) ? akk : a[(long long)i * lda + k];
      lcol[i] = v * rp;
    }
    __syncthreads();

    // ---- rank-1 update of the trailing rows (A columns > k and all of B)
    for (int i = k + 1 + wid; i < n; i += NW) {
      const cd l = lcol[i];
      cd* ar = a + (long long)i * lda;
      for (int j = k + 1 + lane; j < n; j += 64) ar[j] = fmsc(ar[j], l, prowA[j]);
      cd* br = bm + (long long)i * ldb;
      for (int j = lane; j < m; j += 64) br[j] = fmsc(br[j], l, prowB[j]);
    }
    __syncthreads();
  }

  // ---- back substitution U X = B'
  for (int k = n - 1; k >= 0; --k) {
    const cd ukk = a[(long long)k * lda + k];
    for (int j = tid; j < m; j += GESV_THREADS) {
      const cd x = divc(bm[(long long)k * ldb + j], ukk);
      bm[(long long)k * ldb + j] = x;
      prowB[j] = x;
    }
    __syncthreads();
    for (int i = wid; i < k; i += NW) {
      const cd u = a[(long long)i * lda + k];
      cd* br = bm + (long long)i * ldb;
      for (int j = lane; j < m; j += 64) br[j] = fmsc(br[j], u, prowB[j]);
    }
    __syncthreads();
  }
  if (tid == 0 && info) info[item] = first_zero;
}

int zgesv(hipStream_t s, int n, int m, Mat A, Mat B, int batch, int* info) {
  if (batch <= 0 || n <= 0) return 0;
  const int path = lu_path(n);
  if (path == 2 && m >= 64 && batch <= 65535 && A.ld == n && A.stride != 0 && B.stride != 0) {
    const char* mixed = getenv("KHX_LU_MIXED");   // mixed-precision solve (zgesv_refined), off by default
    if (mixed && atoi(mixed) != 0) return zgesv_refined(s, n, m, A, B, batch, info);
  }
  if (path == 2 && (batch == 1 || (A.stride != 0 && B.stride != 0))) return zlu_multi(s, n, m, A, B, batch, nullptr, info, false);
  if (path >= 1) return zlu_blocked(s, n, m, A, B, batch, nullptr, info, false);
  KHX_REQUIRE(A.stride != 0 && B.stride != 0 || batch == 1, "zgesv: operands cannot be broadcast");
  const size_t lds = (size_t)(2 * n + m) * sizeof(cd);
  KHX_REQUIRE(lds <= 150 * 1024, "zgesv: system too large for the LDS staging buffers");
  KHX_TRY(ensure_dynamic_lds((const void*)zgesv_kernel, lds));
  ProfScope ps("zgesv_kernel", s);
  zgesv_kernel<<<dim3(batch), dim3(GESV_THREADS), lds, s>>>(n, m, A.p, A.ld, A.stride, B.p, B.ld,
                                                           B.stride, info);
  KHX_LAUNCH_CHECK();
  return 0;
}

// ======================================================================== elementwise
// Grid: x over (rows*cols) chunks, y over batch chunks.
__global__ void zset_identity_kernel(cd* A, int ld, long long sA, int rows, int cols, cd diag,
                                     int batch) {
  const int total = rows * cols;
  for (long long b = blockIdx.y; b < batch; b += gridDim.y) {
    cd* a = A + b * sA;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
      const int i = e / cols, j = e % cols;
      a[(long long)i * ld + j] = (i == j) ? diag : mkc(0.0, 0.0);
    }
  }
}

__global__ void zaxpby_kernel(cd* D, int ldd, long long sD, cd al, const cd* A, int lda, long long sA,
                              cd be, const cd* B, int ldb, long long sB, int rows, int cols,
                              int batch) {
  const int total = rows * cols;
  for (long long b = blockIdx.y; b < batch; b += gridDim.y) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
      const int i = e / cols, j = e % cols;
      cd v = al * A[b * sA + (long long)i * lda + j];
      if (B) v = fmac(v, be, B[b * sB + (long long)i * ldb + j]);
      D[b * sD + (long long)i * ldd + j] = v;
    }
  }
}

__global__ void zrow_scale_kernel(cd* D, int ldd, long long sD, const cd* d, long long sd,
                                  const cd* A, int lda, long long sA, int rows, int cols, int batch) {
  const int total = rows * cols;
  for (long long b = blockIdx.y; b < batch; b += gridDim.y) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
      const int i = e / cols, j = e % cols;
      D[b * sD + (long long)i * ldd + j] = d[b * sd + i] * A[b * sA + (long long)i * lda + j];
    }
  }
}

static dim3 ew_grid(long long total, int batch) {
  long long gx = (total + 255) / 256;
  if (gx > 64) gx = 64;
  if (gx < 1) gx = 1;
  int gy = batch < 65535 ? batch : 65535;
  return dim3((unsigned)gx, (unsigned)gy);
}

int zset_identity(hipStream_t s, Mat A, int rows, int cols, cd diag, int batch) {
  if (batch <= 0) return 0;
  zset_identity_kernel<<<ew_grid((long long)rows * cols, batch), 256, 0, s>>>(A.p, A.ld, A.stride, rows,
                                                                            cols, diag, batch);
  KHX_LAUNCH_CHECK();
  return 0;
}

int zcopy(hipStream_t s, Mat D, Mat A, int rows, int cols, int batch) {
  return zaxpby(s, D, mkc(1.0, 0.0), A, mkc(0.0, 0.0), Mat{nullptr, 0, 0}, rows, cols, batch);
}

int zaxpby(hipStream_t s, Mat D, cd a, Mat A, cd b, Mat B, int rows, int cols, int batch) {
  if (batch <= 0) return 0;
  zaxpby_kernel<<<ew_grid((long long)rows * cols, batch), 256, 0, s>>>(
      D.p, D.ld, D.stride, a, A.p, A.ld, A.stride, b, B.p, B.ld, B.stride, rows, cols, batch);
  KHX_LAUNCH_CHECK();
  return 0;
}

int zrow_scale(hipStream_t s, Mat D, const cd* d, long long dstride, Mat A, int rows, int cols,
               int batch) {
  if (batch <= 0) return 0;
  zrow_scale_kernel<<<ew_grid((long long)rows * cols, batch), 256, 0, s>>>(
      D.p, D.ld, D.stride, d, dstride, A.p, A.ld, A.stride, rows, cols, batch);
  KHX_LAUNCH_CHECK();
  return 0;
}

}  // namespace khx

namespace khx {

// D[b][i][j] = dl[b][i] * A[b][i][j] * dr[b][j]  (dl / dr may be null = 1)
__global__ void zdiag_scale_kernel(const cd* A, int lda, long long sA, const cd* __restrict__ dl,
                                   const cd* __restrict__ dr, int rows, int cols, int batch, cd* D, int ldd,
                                   long long sD) {
  const long long per = (long long)rows * cols;
  const long long total = per * batch;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / per;
    const int rc = (int)(e % per);
    const int i = rc / cols, j = rc % cols;
    cd v = A[b * sA + (long long)i * lda + j];
    if (dl) v = dl[b * rows + i] * v;
    if (dr) v = v * dr[b * cols + j];
    D[b * sD + (long long)i * ldd + j] = v;
  }
}

// D = I - A
__global__ void zidentity_minus_kernel(const cd* A, int n, int batch, cd* D) {
  const long long per = (long long)n * n;
  const long long total = per * batch;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int rc = (int)(e % per);
    const cd v = A[e];
    D[e] = (rc / n == rc % n) ? mkc(1.0 - v.x, -v.y) : mkc(-v.x, -v.y);
  }
}

static dim3 grid_big(long long total) {
  long long g = (total + 255) / 256;
  if (g > 131072) g = 131072;
  if (g < 1) g = 1;
  return dim3((unsigned)g);
}

int zdiag_scale(hipStream_t s, Mat A, const cd* dl, const cd* dr, int rows, int cols, int batch, Mat D) {
  if (batch <= 0) return 0;
  zdiag_scale_kernel<<<grid_big((long long)rows * cols * batch), 256, 0, s>>>(A.p, A.ld, A.stride, dl, dr, rows,
                                                                              cols, batch, D.p, D.ld, D.stride);
  KHX_LAUNCH_CHECK();
  return 0;
}

int zidentity_minus(hipStream_t s, const cd* A, int n, int batch, cd* D) {
  if (batch <= 0) return 0;
  zidentity_minus_kernel<<<grid_big((long long)n * n * batch), 256, 0, s>>>(A, n, batch, D);
  KHX_LAUNCH_CHECK();
  return 0;
}

}  // namespace khx
