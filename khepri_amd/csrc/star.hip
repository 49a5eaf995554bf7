// Redheffer star product of batched S-matrices (khepri/alternative.py:19-30):
//   D = I - SB11 SA22,  F = I - SA22 SB11
//   S11 = SA11 + SA12 D^-1 SB11 SA21     S12 = SA12 D^-1 SB12
//   S21 = SB21 F^-1 SA21                 S22 = SB22 + SB21 F^-1 SA22 SB12
// Two LU solves with 2n right-hand sides ([SB11 | SB12] and [SA21 | SA22])
// and eight GEMMs per product, all batched.  S must not alias SA or SB.
#include "khx_internal.h"

namespace khx {

__global__ void star_info_kernel(const int* __restrict__ infos, int batch, int* out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < batch) out[b] = infos[b] ? infos[b] : infos[batch + b];
}

size_t star_ws_bytes(int n, int batch) {
  return (size_t)8 * n * n * sizeof(cd) * batch + (size_t)2 * batch * sizeof(int) + 256;
}

int star(hipStream_t s, const cd* SA, long long sA, const cd* SB, long long sB, int n, int batch, cd* S,
         int* info, void* wsv) {
  if (batch <= 0) return 0;
  const long long nn = (long long)n * n;
  const cd one = mkc(1.0, 0.0), mone = mkc(-1.0, 0.0), zero = mkc(0.0, 0.0);
  cd* w = static_cast<cd*>(wsv);
  cd* D = w;
  cd* F = D + nn * batch;
  cd* RD = F + nn * batch;        // n x 2n, ld 2n
  cd* RF = RD + 2 * nn * batch;   // n x 2n
  cd* tmp = RF + 2 * nn * batch;  // n x n
  int* infos = reinterpret_cast<int*>(tmp + nn * batch);
  auto blk = [&](const cd* base, long long str, int q) { return Mat{const_cast<cd*>(base) + q * nn, n, str}; };
  const Mat A11 = blk(SA, sA, 0), A12 = blk(SA, sA, 1), A21 = blk(SA, sA, 2), A22 = blk(SA, sA, 3);
  const Mat B11 = blk(SB, sB, 0), B12 = blk(SB, sB, 1), B21 = blk(SB, sB, 2), B22 = blk(SB, sB, 3);
  const long long sS = 4 * nn;
  const Mat S11 = blk(S, sS, 0), S12 = blk(S, sS, 1), S21 = blk(S, sS, 2), S22 = blk(S, sS, 3);
  const Mat mD{D, n, nn}, mF{F, n, nn}, mT{tmp, n, nn};
  const Mat mRD{RD, 2 * n, 2 * nn}, mRF{RF, 2 * n, 2 * nn};
  const Mat mRD2{RD + n, 2 * n, 2 * nn}, mRF2{RF + n, 2 * n, 2 * nn};

  // D = I - SB11 SA22 ; F = I - SA22 SB11
  KHX_TRY(zset_identity(s, mD, n, n, one, batch));
  KHX_TRY(zgemm(s, n, n, n, mone, B11, A22, one, mD, batch));
  KHX_TRY(zset_identity(s, mF, n, n, one, batch));
  KHX_TRY(zgemm(s, n, n, n, mone, A22, B11, one, mF, batch));
  // [X1 | X2] = D^-1 [SB11 | SB12] ; [Y1 | Y2] = F^-1 [SA21 | SA22]
  KHX_TRY(zcopy(s, mRD, B11, n, n, batch));
  KHX_TRY(zcopy(s, mRD2, B12, n, n, batch));
  KHX_TRY(zcopy(s, mRF, A21, n, n, batch));
  KHX_TRY(zcopy(s, mRF2, A22, n, n, batch));
  KHX_TRY(zgesv(s, n, 2 * n, mD, mRD, batch, infos));
  KHX_TRY(zgesv(s, n, 2 * n, mF, mRF, batch, infos + batch));
  // S11 = SA11 + (SA12 X1) SA21
  KHX_TRY(zgemm(s, n, n, n, one, A12, mRD, zero, mT, batch));
  KHX_TRY(zcopy(s, S11, A11, n, n, batch));
  KHX_TRY(zgemm(s, n, n, n, one, mT, A21, one, S11, batch));
  // S12 = SA12 X2
  KHX_TRY(zgemm(s, n, n, n, one, A12, mRD2, zero, S12, batch));
  // S21 = SB21 Y1
  KHX_TRY(zgemm(s, n, n, n, one, B21, mRF, zero, S21, batch));
  // S22 = SB22 + (SB21 Y2) SB12
  KHX_TRY(zgemm(s, n, n, n, one, B21, mRF2, zero, mT, batch));
  KHX_TRY(zcopy(s, S22, B22, n, n, batch));
  KHX_TRY(zgemm(s, n, n, n, one, mT, B12, one, S22, batch));
  if (info) {
    star_info_kernel<<<dim3((batch + 255) / 256), 256, 0, s>>>(infos, batch, info);
    KHX_LAUNCH_CHECK();
  }
  return 0;
}

}  // namespace khx
