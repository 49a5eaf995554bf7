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

// ------------------------------------------------------------------ adjoint
// SAbar, SBbar (batch, 2, 2, n, n) from Sbar for S = SA * SB (operands not
// broadcast).  D, F and [X1 | X2] = D^-1 [SB11 | SB12], [Y1 | Y2] = F^-1 [SA21 | SA22]
// are recomputed; with S11 = A11 + A12 X1 A21, S12 = A12 X2, S21 = B21 Y1,
// S22 = B22 + B21 Y2 B12 (A = SA, B = SB) the adjoints are
//   A11bar = S11bar, A12bar = S11bar (X1 A21)^H + S12bar X2^H, ...
//   [B11bar | B12bar] += D^-H [X1bar | X2bar], Dbar = -D^-H [X1bar | X2bar] [X1 | X2]^H,
//   and D = I - B11 A22, F = I - A22 B11 pass Dbar, Fbar on to B11 and A22.
size_t star_bwd_ws_bytes(int n, int batch) {
  return (size_t)16 * n * n * sizeof(cd) * batch + (size_t)4 * batch * sizeof(int) + 256;
}

int star_bwd(hipStream_t s, const cd* SA, long long sA, const cd* SB, long long sB, int n, int batch, const cd* Sbar,
             cd* SAbar, cd* SBbar, int* info, void* wsv) {
  if (batch <= 0) return 0;
  const long long nn = (long long)n * n, n2 = 2 * nn, s4 = 4 * nn;
  const cd one = mkc(1.0, 0.0), mone = mkc(-1.0, 0.0), zero = mkc(0.0, 0.0);
  cd* w = static_cast<cd*>(wsv);
  auto take = [&](long long items) {
    cd* p = w;
    w += items * nn * batch;
    return p;
  };
  cd* D = take(1);
  cd* F = take(1);
  cd* DH = take(1);
  cd* FH = take(1);
  cd* RD = take(2);
  cd* RF = take(2);
  cd* ZX = take(2);
  cd* ZY = take(2);
  cd* Db = take(1);
  cd* Fb = take(1);
  cd* T1 = take(1);
  cd* T2 = take(1);
  int* infos = reinterpret_cast<int*>(w);
  auto blk = [&](const cd* base, long long str, int q) { return Mat{const_cast<cd*>(base) + q * nn, n, str}; };
  const Mat A11 = blk(SA, sA, 0), A12 = blk(SA, sA, 1), A21 = blk(SA, sA, 2), A22 = blk(SA, sA, 3);
  const Mat B11 = blk(SB, sB, 0), B12 = blk(SB, sB, 1), B21 = blk(SB, sB, 2), B22 = blk(SB, sB, 3);
  const Mat G11 = blk(Sbar, s4, 0), G12 = blk(Sbar, s4, 1), G21 = blk(Sbar, s4, 2), G22 = blk(Sbar, s4, 3);
  const Mat a11 = blk(SAbar, s4, 0), a12 = blk(SAbar, s4, 1), a21 = blk(SAbar, s4, 2), a22 = blk(SAbar, s4, 3);
  const Mat b11 = blk(SBbar, s4, 0), b12 = blk(SBbar, s4, 1), b21 = blk(SBbar, s4, 2), b22 = blk(SBbar, s4, 3);
  const Mat mD{D, n, nn}, mF{F, n, nn}, mDH{DH, n, nn}, mFH{FH, n, nn}, mDb{Db, n, nn}, mFb{Fb, n, nn};
  const Mat mRD{RD, 2 * n, n2}, mRF{RF, 2 * n, n2}, mX1{RD, 2 * n, n2}, mX2{RD + n, 2 * n, n2};
  const Mat mY1{RF, 2 * n, n2}, mY2{RF + n, 2 * n, n2};
  const Mat mZX{ZX, 2 * n, n2}, mZX2{ZX + n, 2 * n, n2}, mZY{ZY, 2 * n, n2}, mZY2{ZY + n, 2 * n, n2};
  const Mat mT1{T1, n, nn}, mT2{T2, n, nn};

  // forward recomputation
  KHX_TRY(zset_identity(s, mD, n, n, one, batch));
  KHX_TRY(zgemm(s, n, n, n, mone, B11, A22, one, mD, batch));
  KHX_TRY(zset_identity(s, mF, n, n, one, batch));
  KHX_TRY(zgemm(s, n, n, n, mone, A22, B11, one, mF, batch));
  KHX_TRY(conj_transpose(s, mD, n, n, mDH, batch));
  KHX_TRY(conj_transpose(s, mF, n, n, mFH, batch));
  KHX_TRY(zcopy(s, mRD, B11, n, n, batch));
  KHX_TRY(zcopy(s, mX2, B12, n, n, batch));
  KHX_TRY(zcopy(s, mRF, A21, n, n, batch));
  KHX_TRY(zcopy(s, mY2, A22, n, n, batch));
  KHX_TRY(zgesv(s, n, 2 * n, mD, mRD, batch, infos));
  KHX_TRY(zgesv(s, n, 2 * n, mF, mRF, batch, infos + batch));

  // S11 = A11 + A12 X1 A21, S12 = A12 X2
  KHX_TRY(zcopy(s, a11, G11, n, n, batch));
  KHX_TRY(zgemm(s, n, n, n, one, mX1, A21, zero, mT1, batch));
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, one, G11, mT1, zero, a12, batch));
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, one, G12, mX2, one, a12, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, one, A12, G11, zero, mT2, batch));
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, one, mT2, A21, zero, mZX, batch));
  KHX_TRY(zgemm(s, n, n, n, one, A12, mX1, zero, mT1, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, one, mT1, G11, zero, a21, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, one, A12, G12, zero, mZX2, batch));
  // S21 = B21 Y1, S22 = B22 + B21 Y2 B12
  KHX_TRY(zgemm(s, n, n, n, one, mY2, B12, zero, mT1, batch));
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, one, G21, mY1, zero, b21, batch));
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, one, G22, mT1, one, b21, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, one, B21, G21, zero, mZY, batch));
  KHX_TRY(zcopy(s, b22, G22, n, n, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, one, B21, G22, zero, mT2, batch));
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, one, mT2, B12, zero, mZY2, batch));
  KHX_TRY(zgemm(s, n, n, n, one, B21, mY2, zero, mT1, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, one, mT1, G22, zero, b12, batch));
  // [X1 | X2] = D^-1 [B11 | B12]
  KHX_TRY(zgesv(s, n, 2 * n, mDH, mZX, batch, infos + 2 * batch));
  KHX_TRY(zcopy(s, b11, mZX, n, n, batch));
  KHX_TRY(zaxpby(s, b12, one, b12, one, mZX2, n, n, batch));
  KHX_TRY(zgemm_op(s, 0, 2, n, n, 2 * n, mone, mZX, mRD, zero, mDb, batch));
  // [Y1 | Y2] = F^-1 [A21 | A22]
  KHX_TRY(zgesv(s, n, 2 * n, mFH, mZY, batch, infos + 3 * batch));
  KHX_TRY(zaxpby(s, a21, one, a21, one, mZY, n, n, batch));
  KHX_TRY(zcopy(s, a22, mZY2, n, n, batch));
  KHX_TRY(zgemm_op(s, 0, 2, n, n, 2 * n, mone, mZY, mRF, zero, mFb, batch));
  // D = I - B11 A22, F = I - A22 B11
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, mone, mDb, A22, one, b11, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, mone, B11, mDb, one, a22, batch));
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, mone, mFb, B11, one, a22, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, mone, A22, mFb, one, b11, batch));
  if (info) {
    star_info_kernel<<<dim3((batch + 255) / 256), 256, 0, s>>>(infos, batch, info);
    KHX_LAUNCH_CHECK();
  }
  return 0;
}

}  // namespace khx
