// Dense linear algebra of the engine (FP64): a batched GEMM on the CDNA matrix cores
// (v_mfma_f64_16x16x4f64) and a cyclic two-sided Jacobi eigensolver with one workgroup per
// symmetric matrix (parallel tournament ordering, the matrix in LDS when it fits).
#include "common.hpp"

namespace axtb {

// C = alpha op(A) diag(dk) op(B) + beta C, column-major; op(A) is m x k, op(B) k x n.
struct GemmJob {
  int m, n, k, ta, tb;
  const double* A;
  const double* B;
  double* C;
  const double* dk;  // optional scaling of the inner index
  long long lda, ldb, ldc;
  double alpha, beta;
  // optional second term (128-tile kernel only): C = alpha (op(A) op(B) + op(A2) op(B2)) + beta C,
  // A2, B2 with the leading dimensions and transposes of A, B and inner size k2 (no scaling)
  const double* A2 = nullptr;
  const double* B2 = nullptr;
  int k2 = 0;
  // 128-tile kernel only: compute only the tiles that hold elements of the lower triangle (row >=
  // column) of C
  int lower = 0;
  // the refinement update C = A + A (B - I) with the product on the FP32 matrix cores (A, B square
  // n x n, no transposes; k_gemm128_upd32)
  int f32upd = 0;
  // optional device flag: the job runs only while *run != 0 (set and cleared by the device, so that
  // a launch over a fixed job list needs no host decision)
  const int* run = nullptr;
};

__device__ __forceinline__ bool job_skipped(const int* run) { return run && *run == 0; }

typedef double d4v __attribute__((ext_vector_type(4)));

// Load through the global address space: the GEMM operands' pointers come from job structs in
// memory and would otherwise be generic (flat) loads, which the LDS waits of the MFMA loop also
// wait for.
typedef const double __attribute__((address_space(1)))* gdp;
__device__ __forceinline__ double gload(const double* p) { return *(gdp)p; }

constexpr int GT = 64;   // tile edge
constexpr int GK = 16;   // k step

// 256 threads = 4 wavefronts; wavefront w owns the 32 x 32 quadrant (w & 1, w >> 1) of the tile
// as 2 x 2 MFMA blocks of 16 x 16. LDS holds op(A) as As[k][row] and op(B) as Bs[k][col].
__global__ __launch_bounds__(256) void k_gemm(const GemmJob* jobs) {
  const GemmJob J = jobs[blockIdx.z];
  const int row0 = blockIdx.y * GT, col0 = blockIdx.x * GT;
  if (row0 >= J.m || col0 >= J.n || job_skipped(J.run)) return;
  __shared__ double As[2][GK][GT + 1];
  __shared__ double Bs[2][GK][GT + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = (w & 1) * 32, wc = (w >> 1) * 32;
  d4v acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = (d4v){0.0, 0.0, 0.0, 0.0};
  // element q (of 4) of this thread's share of a k-step of op(A) (row r, k) and op(B) (col c, k);
  // loads read clamped in-bounds addresses, masks and the inner scaling are applied at the store
  // to LDS (the next step's loads are in flight during this step's products)
  auto a_rk = [&](int q, int& r, int& k) {
    const int e = tid + 256 * q;
    if (J.ta == 0) {
      r = e % GT;
      k = e / GT;
    } else {
      k = e % GK;
      r = e / GK;
    }
  };
  auto b_ck = [&](int q, int& c, int& k) {
    const int e = tid + 256 * q;
    if (J.tb == 0) {
      k = e % GK;
      c = e / GK;
    } else {
      c = e % GT;
      k = e / GT;
    }
  };
  double ra[4], rb[4], rd[4];
  unsigned ma = 0u, mb = 0u;
  auto load = [&](int k0) {
    ma = mb = 0u;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int r, k;
      a_rk(q, r, k);
      const int gr = row0 + r, gk = k0 + k;
      const int grc = min(gr, J.m - 1), gkc = min(gk, J.k - 1);
      ra[q] = gload(J.ta == 0 ? J.A + grc + (long long)gkc * J.lda : J.A + gkc + (long long)grc * J.lda);
      ma |= (gr < J.m && gk < J.k) ? 1u << q : 0u;
      int c;
      b_ck(q, c, k);
      const int gc = col0 + c, gkb = k0 + k;
      const int gcc = min(gc, J.n - 1), gkbc = min(gkb, J.k - 1);
      rb[q] = gload(J.tb == 0 ? J.B + gkbc + (long long)gcc * J.ldb : J.B + gcc + (long long)gkbc * J.ldb);
      mb |= (gc < J.n && gkb < J.k) ? 1u << q : 0u;
    }
    if (J.dk) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int r, k;
        a_rk(q, r, k);
        rd[q] = gload(J.dk + min(k0 + k, J.k - 1));
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int r, k;
      a_rk(q, r, k);
      const double av = J.dk ? ra[q] * rd[q] : ra[q];
      As[buf][k][r] = (ma >> q) & 1u ? av : 0.0;
      int c;
      b_ck(q, c, k);
      Bs[buf][k][c] = (mb >> q) & 1u ? rb[q] : 0.0;
    }
  };
  const int nk = (J.k + GK - 1) / GK;
  if (nk > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * GK);
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      const int kl = kk + (lane >> 4), li = lane & 15;
      const double a0 = As[buf][kl][wr + li], a1 = As[buf][kl][wr + 16 + li];
      const double b0 = Bs[buf][kl][wc + li], b1 = Bs[buf][kl][wc + 16 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    if (kt + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }
  // lane l, element r of a 16 x 16 block holds row l / 16 + 4 r, column l % 16 (operands: lane l
  // holds A(l % 16, k0 + l / 16) and B(k0 + l / 16, l % 16))
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int r = 0; r < 4; r++) {
        int gr = row0 + wr + 16 * i + (lane >> 4) + 4 * r;
        int gc = col0 + wc + 16 * j + (lane & 15);
        if (gr < J.m && gc < J.n) {
          double* c = &J.C[gr + (long long)gc * J.ldc];
          double v = J.alpha * acc[i][j][r];
          *c = J.beta == 0.0 ? v : v + J.beta * *c;
        }
      }
}

// Large-tile variant: 128 x 128 output tile per workgroup of 4 wavefronts (each 64 x 64 as 4 x 4
// MFMA blocks), K in steps of 16 through double-buffered LDS with the next step's global loads
// held in registers while the current one is multiplied; one barrier per step. TA / TB select
// op(A) = A^T / op(B) = B^T at compile time. Every job of a launch has the same TA, TB.
constexpr int BT = 128;
constexpr int BK = 16;
constexpr int BP = BT + 1;  // padded LDS row (an odd row length measured fastest)

template <int TA, int TB, bool TWO>
__global__ __launch_bounds__(256, 2) void k_gemm128(const GemmJob* jobs) {
  const GemmJob J = jobs[blockIdx.z];
  const int row0 = blockIdx.y * BT, col0 = blockIdx.x * BT;
  if (row0 >= J.m || col0 >= J.n || job_skipped(J.run)) return;
  if (J.lower && col0 > row0 + BT - 1) return;
  __shared__ double As[2][BK][BP];
  __shared__ double Bs[2][BK][BP];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = (w & 1) * 64, wc = (w >> 1) * 64;
  double ra[8], rb[8];
  // element q of this thread's share of a k-step: (tile row / col, k)
  auto a_rk = [&](int q, int& r, int& k) {
    if (TA == 0) {
      r = tid & 127;
      k = (tid >> 7) * 8 + q;
    } else {
      k = tid & 15;
      r = (tid >> 4) + 16 * q;
    }
  };
  auto b_ck = [&](int q, int& c, int& k) {
    if (TB == 1) {
      c = tid & 127;
      k = (tid >> 7) * 8 + q;
    } else {
      k = tid & 15;
      c = (tid >> 4) + 16 * q;
    }
  };
  // inner index gk < k: first term; k <= gk < k + k2: second term at gk - k. Branch-free: every
  // load reads a clamped (in-bounds) address; the masks ma / mb (bit q: element q inside op(A),
  // op(B)) and the inner scaling dk of the first term are applied when the step is stored to LDS,
  // so nothing waits for the loads before the products of the current step.
  const int kt = J.k + (TWO ? J.k2 : 0);
  double rd[8];
  unsigned ma = 0u, mb = 0u;
  auto load = [&](int k0) {
    ma = mb = 0u;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int r, k;
      a_rk(q, r, k);
      const int gr = row0 + r, gk = k0 + k;
      const int grc = min(gr, J.m - 1);
      const bool first = !TWO || gk < J.k || J.k2 <= 0;
      const double* pa = first ? J.A : J.A2;
      const int gkc = first ? min(gk, J.k - 1) : min(gk - J.k, J.k2 - 1);
      ra[q] = gload(TA == 0 ? pa + grc + (long long)gkc * J.lda : pa + gkc + (long long)grc * J.lda);
      ma |= (gr < J.m && gk < kt) ? 1u << q : 0u;
      int c;
      b_ck(q, c, k);
      const int gc = col0 + c, gkb = k0 + k;
      const int gcc = min(gc, J.n - 1);
      const bool firstb = !TWO || gkb < J.k || J.k2 <= 0;
      const double* pb = firstb ? J.B : J.B2;
      const int gkbc = firstb ? min(gkb, J.k - 1) : min(gkb - J.k, J.k2 - 1);
      rb[q] = gload(TB == 0 ? pb + gkbc + (long long)gcc * J.ldb : pb + gcc + (long long)gkbc * J.ldb);
      mb |= (gc < J.n && gkb < kt) ? 1u << q : 0u;
    }
    if (!TWO && J.dk) {  // two-term products carry no inner scaling
#pragma unroll
      for (int q = 0; q < 8; q++) {
        int r, k;
        a_rk(q, r, k);
        rd[q] = gload(J.dk + min(k0 + k, J.k - 1));
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int r, k;
      a_rk(q, r, k);
      const double av = (!TWO && J.dk) ? ra[q] * rd[q] : ra[q];
      As[buf][k][r] = (ma >> q) & 1u ? av : 0.0;
      int c;
      b_ck(q, c, k);
      Bs[buf][k][c] = (mb >> q) & 1u ? rb[q] : 0.0;
    }
  };
  d4v acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = (d4v){0.0, 0.0, 0.0, 0.0};
  const int nk = (kt + BK - 1) / BK;
  if (nk > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kl = kk + (lane >> 4), li = lane & 15;
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        a[i] = As[buf][kl][wr + 16 * i + li];
        b[i] = Bs[buf][kl][wc + 16 * i + li];
      }
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }
  // lane l, element r of a 16 x 16 block holds row l / 16 + 4 r, column l % 16
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gr = row0 + wr + 16 * i + (lane >> 4) + 4 * r;
        const int gc = col0 + wc + 16 * j + (lane & 15);
        if (gr < J.m && gc < J.n) {
          double* c = &J.C[gr + (long long)gc * J.ldc];
          const double v = J.alpha * acc[i][j][r];
          *c = J.beta == 0.0 ? v : v + J.beta * *c;
        }
      }
}

// The update of a refinement step, C = U + U E with E = B - I (B holds I + E): the product U E,
// whose entries are the size of the correction, runs in FP32 (operands rounded to FP32 when they
// are staged, FP32 accumulation on the matrix cores, v_mfma_f32_16x16x4f32); the sum with U is
// FP64. The rounding of the product is |E| 2^-24 sqrt(n) relative to U and is itself corrected by
// the next refinement step, whose residuals (U^T A U, U^T U) stay FP64. 128 x 128 tiles, 4
// wavefronts of 64 x 64, double-buffered LDS as in k_gemm128.
typedef float f4v __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256, 2) void k_gemm128_upd32(const GemmJob* jobs) {
  const GemmJob J = jobs[blockIdx.z];
  const int row0 = blockIdx.y * BT, col0 = blockIdx.x * BT;
  if (row0 >= J.m || col0 >= J.n || job_skipped(J.run)) return;
  __shared__ float As[2][BK][BP];
  __shared__ float Bs[2][BK][BP];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = (w & 1) * 64, wc = (w >> 1) * 64;
  double ra[8], rb[8];
  unsigned ma = 0u, mb = 0u;
  // element q of this thread's share of a k-step: A (no transpose) row r = tid & 127, k = (tid >> 7) * 8 + q;
  // B (no transpose) k = tid & 15, column c = (tid >> 4) + 16 q
  auto load = [&](int k0) {
    ma = mb = 0u;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int r = tid & 127, k = (tid >> 7) * 8 + q;
      const int gr = row0 + r, gk = k0 + k;
      ra[q] = gload(J.A + min(gr, J.m - 1) + (long long)min(gk, J.k - 1) * J.lda);
      ma |= (gr < J.m && gk < J.k) ? 1u << q : 0u;
      const int kb = tid & 15, c = (tid >> 4) + 16 * q;
      const int gc = col0 + c, gkb = k0 + kb;
      rb[q] = gload(J.B + min(gkb, J.k - 1) + (long long)min(gc, J.n - 1) * J.ldb) - (gkb == gc ? 1.0 : 0.0);
      mb |= (gc < J.n && gkb < J.k) ? 1u << q : 0u;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      As[buf][(tid >> 7) * 8 + q][tid & 127] = (ma >> q) & 1u ? (float)ra[q] : 0.0f;
      Bs[buf][tid & 15][(tid >> 4) + 16 * q] = (mb >> q) & 1u ? (float)rb[q] : 0.0f;
    }
  };
  f4v acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = (f4v){0.0f, 0.0f, 0.0f, 0.0f};
  const int nk = (J.k + BK - 1) / BK;
  if (nk > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kl = kk + (lane >> 4), li = lane & 15;
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        a[i] = As[buf][kl][wr + 16 * i + li];
        b[i] = Bs[buf][kl][wc + 16 * i + li];
      }
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }
  // lane l, element r of a 16 x 16 FP32 block holds row 4 (l / 16) + r, column l % 16
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gr = row0 + wr + 16 * i + 4 * (lane >> 4) + r;
        const int gc = col0 + wc + 16 * j + (lane & 15);
        if (gr < J.m && gc < J.n) J.C[gr + (long long)gc * J.ldc] = J.A[gr + (long long)gc * J.lda] + (double)acc[i][j][r];
      }
}

// The operands of the FP32 update by a library product: X (m x n, leading dimension ld) rounded to
// FP32 into Y (leading dimension m), minus the identity when eye is set (B = I + E -> E).
__global__ void k_to_f32(const double* X, long long ld, int m, int n, int eye, float* Y) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)m * n) return;
  const int r = (int)(e % m), c = (int)(e / m);
  Y[e] = (float)(X[r + c * ld] - (eye && r == c ? 1.0 : 0.0));
}

// C = U + P: the FP32 product P (m x n, leading dimension m) added to U in FP64.
__global__ void k_add_f32(const double* U, long long ldu, const float* P, int m, int n, double* C, long long ldc) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)m * n) return;
  const int r = (int)(e % m), c = (int)(e / m);
  C[r + c * ldc] = U[r + c * ldu] + (double)P[e];
}

// C = sum over s of the split-K partials P_s (m x n, ld m, back to back) + beta C, the partials
// in a fixed order.
__global__ void k_gemm_reduce(const double* P, int splits, int m, int n, double* C, long long ldc, double beta) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)m * n) return;
  const int r = (int)(e % m), c = (int)(e / m);
  double s = 0.0;
  for (int q = 0; q < splits; q++) s += P[(long long)q * m * n + e];
  double* out = &C[r + c * ldc];
  *out = beta == 0.0 ? s : s + beta * *out;
}

// Batched products of small matrices (m, n <= 128): one workgroup of 8 wavefronts per job. The
// 8 x 8 grid of 16 x 16 MFMA output blocks is split 4 x 2 over the wavefronts: wavefront
// (wr, wc) owns block rows wr and wr + 4 and block columns wc, wc + 2, wc + 4, wc + 6 (blocks past
// the matrix edge skipped), so each k sub-step reads 2 A and 4 B fragments from LDS for up to 8
// MFMAs. K in steps of 16 through double-buffered LDS with the next step's loads in registers
// (clamped addresses, masks applied at the store, as in k_gemm128). Few registers per thread: two
// workgroups share a CU.
constexpr int SP = 129;  // padded LDS row of the small-matrix kernel (odd, as BP)
template <int TA, int TB>
__global__ __launch_bounds__(512) void k_gemm_small(const GemmJob* jobs) {
  const GemmJob J = jobs[blockIdx.x];
  if (job_skipped(J.run)) return;
  __shared__ double As[2][BK][SP];
  __shared__ double Bs[2][BK][SP];
  // the wavefront index as a scalar: the block branches are then uniform
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = w & 3, wc = w >> 2;
  const int mb = (J.m + 15) >> 4, nbk = (J.n + 15) >> 4;
  // lower (square C): only the blocks on or below the diagonal, the others written as their mirror
  const bool low = J.lower && J.m == J.n && J.beta == 0.0;
  auto live = [&](int i, int j) {
    const int bi = wr + 4 * i, bj = wc + 2 * j;
    return bi < mb && bj < nbk && (!low || bi >= bj);
  };
  // element q (of 4) of this thread's share of a k-step: (row / col within 128, k within 16)
  auto a_rk = [&](int q, int& r, int& k) {
    if (TA == 0) {
      r = tid & 127;
      k = (tid >> 7) * 4 + q;
    } else {
      k = tid & 15;
      r = (tid >> 4) + 32 * q;
    }
  };
  auto b_ck = [&](int q, int& c, int& k) {
    if (TB == 1) {
      c = tid & 127;
      k = (tid >> 7) * 4 + q;
    } else {
      k = tid & 15;
      c = (tid >> 4) + 32 * q;
    }
  };
  double ra[4], rb[4], rd[4];
  unsigned ma = 0u, mbk = 0u;
  auto load = [&](int k0) {
    ma = mbk = 0u;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int r, k;
      a_rk(q, r, k);
      const int gk = k0 + k, grc = min(r, J.m - 1), gkc = min(gk, J.k - 1);
      ra[q] = gload(TA == 0 ? J.A + grc + (long long)gkc * J.lda : J.A + gkc + (long long)grc * J.lda);
      ma |= (r < J.m && gk < J.k) ? 1u << q : 0u;
      int c;
      b_ck(q, c, k);
      const int gkb = k0 + k, gcc = min(c, J.n - 1), gkbc = min(gkb, J.k - 1);
      rb[q] = gload(TB == 0 ? J.B + gkbc + (long long)gcc * J.ldb : J.B + gcc + (long long)gkbc * J.ldb);
      mbk |= (c < J.n && gkb < J.k) ? 1u << q : 0u;
    }
    if (J.dk) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int r, k;
        a_rk(q, r, k);
        rd[q] = gload(J.dk + min(k0 + k, J.k - 1));
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int r, k;
      a_rk(q, r, k);
      const double av = J.dk ? ra[q] * rd[q] : ra[q];
      As[buf][k][r] = (ma >> q) & 1u ? av : 0.0;
      int c;
      b_ck(q, c, k);
      Bs[buf][k][c] = (mbk >> q) & 1u ? rb[q] : 0.0;
    }
  };
  d4v acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = (d4v){0.0, 0.0, 0.0, 0.0};
  const int nk = (J.k + BK - 1) / BK;
  if (nk > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kl = kk + (lane >> 4), li = lane & 15;
      // rows 16 (wr + 4 i) + li and columns 16 (wc + 2 j) + li stay below SP: no bounds test
      double a[2], b[4];
#pragma unroll
      for (int i = 0; i < 2; i++) a[i] = As[buf][kl][16 * (wr + 4 * i) + li];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = Bs[buf][kl][16 * (wc + 2 * j) + li];
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (live(i, j)) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (!live(i, j)) continue;
      const int bi = wr + 4 * i, bj = wc + 2 * j;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gr = 16 * bi + (lane >> 4) + 4 * r, gc = 16 * bj + (lane & 15);
        if (gr < J.m && gc < J.n) {
          double* c = &J.C[gr + (long long)gc * J.ldc];
          const double v = J.alpha * acc[i][j][r];
          const double o = J.beta == 0.0 ? v : v + J.beta * *c;
          *c = o;
          if (low && bi != bj) J.C[gc + (long long)gr * J.ldc] = o;  // the mirror above the diagonal
        }
      }
    }
}

static constexpr auto k_gemm_small_nn = k_gemm_small<0, 0>;
static constexpr auto k_gemm_small_nt = k_gemm_small<0, 1>;
static constexpr auto k_gemm_small_tn = k_gemm_small<1, 0>;
static constexpr auto k_gemm_small_tt = k_gemm_small<1, 1>;

static constexpr auto k_gemm128_nn = k_gemm128<0, 0, false>;
static constexpr auto k_gemm128_nt = k_gemm128<0, 1, false>;
static constexpr auto k_gemm128_tn = k_gemm128<1, 0, false>;
static constexpr auto k_gemm128_tt = k_gemm128<1, 1, false>;
static constexpr auto k_gemm128_nt2 = k_gemm128<0, 1, true>;  // C = alpha (A B^T + A2 B2^T) + beta C

// ------------------------------------------------------------------ Jacobi eigensolver
struct EigJob {
  int n;
  double* A;      // n x n (ld n), destroyed
  double* V;      // n x n work (ld n)
  double* w;      // eigenvalues, ascending
  double* Z;      // eigenvectors (ld n), columns in the order of w; null: eigenvalues only (V unused)
  const int* run = nullptr;  // optional device flag: the job runs only while *run != 0
};


// Pair k of round r of the round-robin tournament over nn (even) indices: index 0 stays, the
// others rotate (r < nn - 1, so one conditional subtraction replaces the modulo).
__device__ __forceinline__ void tour_pair(int r, int k, int nn, int& p, int& q) {
  auto pos = [&](int s) {
    if (s == 0) return 0;
    int x = s - 1 + r;
    if (x >= nn - 1) x -= nn - 1;
    return 1 + x;
  };
  p = pos(k);
  q = pos(nn - 1 - k);
  if (p > q) {
    int t = p;
    p = q;
    q = t;
  }
}

// One workgroup per matrix. A round of the tournament rotates n/2 disjoint pairs (p_k, q_k) at
// once; A <- J^T A J then transforms every 2 x 2 block (pair k1 rows, pair k2 columns) on its own,
// B' = R_k1^T B R_k2 with R = [[c, s], [-s, c]], so one pass over the lower blocks (mirrored)
// updates the whole matrix. V <- V J alongside. MODE 0: A and V in global memory; 1: A in LDS;
// 2: A and V in LDS. A sweep starts with a test of every off-diagonal element against the
// rotation threshold, so a converged matrix costs one pass, not a sweep.
// status[job] = number of sweeps (negative if not converged within max_sweeps).
// Rotation thresholds per arithmetic: an element is rotated while |a_pq| > REL sqrt(|a_pp a_qq|)
// and |a_pq| > ABS max|A|; in FP32 the stop is at the rounding level (the result is refined in
// FP64 afterwards).
template <typename T> struct JacobiTol;

template <> struct JacobiTol<double> {
  static constexpr double REL = 1e-15, ABS = 1e-17;
};
template <> struct JacobiTol<float> {
  static constexpr float REL = 2.5e-7f, ABS = 2.5e-7f;
};

// rotation of one pair of a round: one 16-byte LDS record in FP32 (32 bytes in FP64)
template <typename T> struct alignas(16) JRot {
  T c, s;
  int p, q;
};

template <int MODE, typename T, bool PACK>
__global__ __launch_bounds__(1024) void k_jacobi(const EigJob* jobs, int max_sweeps, int* status, float tol_scale) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* lds = (T*)lds_raw;
  __shared__ int rotated;
  __shared__ int rflag[2][2];  // per rotation buffer and rotation wavefront: any pair rotated
  __shared__ T anorm;
  const EigJob J = jobs[blockIdx.x];
  const int n = J.n, tid = threadIdx.x, nt = blockDim.x;
  if (job_skipped(J.run)) {
    if (tid == 0) status[blockIdx.x] = JACOBI_SKIPPED;
    return;
  }
  if (n <= 0) return;
  const int nn = n + (n & 1), np = nn / 2;
  const int lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
  // dynamic LDS: [A (MODE >= 1; PACK: the lower triangle packed by rows, element (r, c), r >= c,
  // at r (r + 1) / 2 + c; else full storage, columns ld = n | 1 apart)] [V (MODE 2; columns ldv
  // apart)] [JRot per pair: 2 buffers, 16-byte aligned]; MODE 0 works on J.A (full storage) and
  // J.V in place (FP64 only). Only the lower triangle of A is kept up to date: element (r, c) is
  // read and written at (max, min). An odd column distance in LDS keeps the column-strided
  // accesses of a round (consecutive threads on consecutive pairs) on distinct banks.
  const int ld = (PACK || MODE == 0) ? n : (n | 1);
  const int ldv = (MODE == 2 && !PACK) ? ld : n;
  const int nl = PACK ? n * (n + 1) / 2 : n * ld;
  T* A = MODE >= 1 ? lds : (T*)J.A;
  T* V = MODE == 2 ? lds + nl : (T*)J.V;
  const int rot_off = ((((MODE >= 1 ? nl : 0) + (MODE == 2 ? n * ldv : 0)) * (int)sizeof(T) + 15) / 16) * 16;
  JRot<T>* rot0 = (JRot<T>*)(lds_raw + rot_off);
  // element (r, c) of the symmetric A
  auto aix = [&](int r, int c) -> int {
    const int hi = max(r, c), lo = min(r, c);
    return PACK ? hi * (hi + 1) / 2 + lo : hi + lo * ld;
  };
  if (MODE >= 1) {
    if (PACK) {
      for (int e = tid; e < n * n; e += nt) {
        const int r = e % n, c = e / n;
        if (r >= c) A[r * (r + 1) / 2 + c] = (T)J.A[e];
      }
    } else {
      for (int e = tid; e < n * ld; e += nt) {  // the padding row (n even) zeroed
        const int r = e % ld, c = e / ld;
        A[e] = r < n ? (T)J.A[r + c * n] : (T)0;
      }
    }
  }
  // eigenvalues only (J.Z null): V is neither formed nor read
  const bool vec = J.Z != nullptr;
  for (int e = tid; vec && e < n * ldv; e += nt) {
    const int r = e % ldv, c = e / ldv;
    V[e] = r == c ? (T)1 : (T)0;
  }
  if (tid == 0) rotated = 0;
  if (tid < 4) rflag[tid >> 1][tid & 1] = 0;
  __syncthreads();
  if (tid < 64) {
    T s = 0;
    for (int e = tid; e < nl; e += 64) s = fmax(s, fabs(A[e]));
    for (int o = 32; o > 0; o >>= 1) s = fmax(s, __shfl_xor(s, o, 64));
    if (tid == 0) anorm = s;
  }
  __syncthreads();
  const T an = anorm;
  const T REL = JacobiTol<T>::REL * (T)tol_scale, ABS = JacobiTol<T>::ABS * (T)tol_scale;
  // the rotation test of a pair (the same at the start of a sweep and in a round)
  auto need = [&](T app, T aqq, T apq) {
    return fabs(apq) > REL * sqrt(fabs(app) * fabs(aqq)) && fabs(apq) > ABS * an;
  };
  const int nblk = np * (np + 1) / 2;
  // this thread's lower 2 x 2 blocks (k1 >= k2; block b = k1 (k1 + 1) / 2 + k2, b = tid + j nt),
  // the same in every round
  constexpr int MAXB = 10;  // nblk <= 5050 for n <= 200
  int bk1[MAXB], bk2[MAXB];
  {
    int k1 = 0, k2 = tid;
    while (k2 > k1) {
      k2 -= k1 + 1;
      k1++;
    }
#pragma unroll
    for (int j = 0; j < MAXB; j++) {
      bk1[j] = k1;
      bk2[j] = k2;
      k2 += nt;
      while (k2 > k1) {
        k2 -= k1 + 1;
        k1++;
      }
    }
  }
  // rotation (c, s) and pair (p, q) of every pair k of round r into buffer `buf`. With V in LDS
  // (MODE 2) the last ceil(np / 64) wavefronts compute them (pair k on thread nt - 1 - k) while
  // the first nvw wavefronts update V; with V in global memory every wavefront updates V and the
  // first threads compute the rotations
  const int nvw = MODE == 2 ? max(1, nw - (np + 63) / 64) : nw;
  const int rt = MODE == 2 ? nt - 1 - tid : tid;
  // Pair k is on thread rt == k (np <= 100 < nt, so pairs k < 64 share one wavefront and the
  // others a second); each wavefront that holds pairs sets rflag[buf][k / 64] to whether any of its
  // pairs rotates, so that a round with no rotation skips its updates. Called by every thread.
  auto rotations = [&](int r, int buf) {
    JRot<T>* rb = rot0 + buf * np;
    const int k = rt;
    bool rotd = false;
    if (k < np) {
      int p, q;
      tour_pair(r, k, nn, p, q);
      T c = 1, s = 0;
      if (q < n) {
        const T app = A[aix(p, p)], aqq = A[aix(q, q)], apq = A[aix(p, q)];
        if (need(app, aqq, apq)) {
          const T tau = (aqq - app) / (2 * apq);
          const T t = (tau >= 0 ? (T)1 : (T)-1) / (fabs(tau) + sqrt(1 + tau * tau));
          c = 1 / sqrt(1 + t * t);
          s = t * c;
        }
      }
      rb[k] = JRot<T>{c, s, p, q};
      rotd = s != 0;
    }
    const bool any = __ballot(rotd) != 0ull;
    if (k < np && (k & 63) == 0) rflag[buf][k >> 6] = any ? 1 : 0;
  };
  int sweep = 0;
  bool conv = n == 1;
  for (sweep = 0; sweep < max_sweeps && !conv; sweep++) {
    // convergence test over all off-diagonal elements (i < j)
    {
      int j = tid / n, i = tid % n;
      const int dj = nt / n, dii = nt % n;
      for (int e = tid; e < n * n; e += nt) {
        if (i < j) {
          const T aij = A[aix(i, j)];
          if (need(A[aix(i, i)], A[aix(j, j)], aij)) rotated = 1;
        }
        i += dii;
        j += dj;
        if (i >= n) {
          i -= n;
          j++;
        }
      }
    }
    __syncthreads();
    conv = rotated == 0;
    __syncthreads();
    if (conv) break;
    if (tid == 0) rotated = 0;
    rotations(0, 0);
    __syncthreads();
    for (int r = 0; r < nn - 1; r++) {
      const int buf = r & 1;
      const JRot<T>* rb = rot0 + buf * np;
      const bool any = (rflag[buf][0] | rflag[buf][1]) != 0;  // a round with no rotation: no updates
      // A <- J^T A J on the lower 2 x 2 blocks of this thread. A block whose pairs both rotate by
      // identity is skipped; the four elements are read unconditionally (an index q >= n, odd n, is
      // clamped to a valid element whose value is never used or written).
#pragma unroll
      for (int jb = 0; jb < MAXB; jb++) {
        if (!any || tid + jb * nt >= nblk) break;
        const JRot<T> r1 = rb[bk1[jb]], r2 = rb[bk2[jb]];
        const T s1 = r1.s, s2 = r2.s;
        if (s1 == 0 && s2 == 0) continue;
        const T c1 = r1.c, c2 = r2.c;
        const int p1 = r1.p, p2 = r2.p;
        const bool vq1 = r1.q < n, vq2 = r2.q < n;
        const int q1 = vq1 ? r1.q : p1, q2 = vq2 ? r2.q : p2;
        const int i00 = aix(p1, p2), i01 = aix(p1, q2), i10 = aix(q1, p2), i11 = aix(q1, q2);
        const T e00 = A[i00], e01 = A[i01], e10 = A[i10], e11 = A[i11];
        const T b00 = e00, b01 = vq2 ? e01 : (T)0, b10 = vq1 ? e10 : (T)0, b11 = (vq1 && vq2) ? e11 : (T)0;
        const T t00 = c2 * b00 - s2 * b01, t01 = s2 * b00 + c2 * b01;
        const T t10 = c2 * b10 - s2 * b11, t11 = s2 * b10 + c2 * b11;
        const T n00 = c1 * t00 - s1 * t10, n01 = c1 * t01 - s1 * t11;
        const T n10 = s1 * t00 + c1 * t10, n11 = s1 * t01 + c1 * t11;
        if (p1 == p2) {  // diagonal block (k1 == k2)
          A[i00] = n00;
          if (vq1) {
            A[i11] = n11;
            A[i01] = 0;
          }
          continue;
        }
        A[i00] = n00;
        if (vq2) A[i01] = n01;
        if (vq1) A[i10] = n10;
        if (vq1 && vq2) A[i11] = n11;
      }
      __syncthreads();
      // the next round's rotations (from the updated A) beside this round's V <- V J
      if (r + 1 < nn - 1) rotations(r + 1, buf ^ 1);
      // V: one pair per wavefront at a time (pairs wv, wv + nvw, ...), the lanes over consecutive
      // rows; a pair that does not rotate costs one record read. With V in LDS the last wavefronts
      // compute the next round's rotations instead.
      for (int k = wv; vec && any && k < np && wv < nvw; k += nvw) {
        const JRot<T> rk = rb[k];
        const T sn = rk.s;
        if (sn == 0) continue;
        const T c = rk.c;
        T* vp = V + rk.p * ldv;
        T* vq = V + rk.q * ldv;
        for (int i = lane; i < n; i += 64) {
          const T a = vp[i], b = vq[i];
          vp[i] = c * a - sn * b;
          vq[i] = sn * a + c * b;
        }
      }
      __syncthreads();
    }
  }
  // eigenvalues ascending (ties by index), vectors permuted alike
  for (int i = tid; i < n; i += nt) {
    const T di = A[aix(i, i)];
    int rank = 0;
    for (int j = 0; j < n; j++) {
      const T dj = A[aix(j, j)];
      if (dj < di || (dj == di && j < i)) rank++;
    }
    J.w[rank] = (double)di;
    for (int r = 0; vec && r < n; r++) J.Z[r + (long long)rank * n] = (double)V[r + (long long)i * ldv];
  }
  if (tid == 0) status[blockIdx.x] = conv ? sweep : -sweep;
}

template __global__ void k_jacobi<0, double, false>(const EigJob*, int, int*, float);
template __global__ void k_jacobi<1, double, true>(const EigJob*, int, int*, float);
template __global__ void k_jacobi<2, double, true>(const EigJob*, int, int*, float);
template __global__ void k_jacobi<1, float, true>(const EigJob*, int, int*, float);
template __global__ void k_jacobi<2, float, false>(const EigJob*, int, int*, float);

using JacobiKernel = void (*)(const EigJob*, int, int*, float);
// the instantiation for (mode, arithmetic); A packed except in the FP32 LDS mode 2
inline JacobiKernel jacobi_kernel(int mode, bool fp32) {
  if (fp32) return mode == 2 ? k_jacobi<2, float, false> : k_jacobi<1, float, true>;
  if (mode == 2) return k_jacobi<2, double, true>;
  if (mode == 1) return k_jacobi<1, double, true>;
  return k_jacobi<0, double, false>;
}

// One Ogita-Aishima refinement step of approximate eigenvectors U of A (FP64), given
// S = U^T A U and G = U^T U (R = I - G): lambda_i = s_ii / g_ii;
//   e_ij = (s_ij - lambda_j g_ij) / (lambda_j - lambda_i)  if |lambda_i - lambda_j| > delta,
//   e_ij = -g_ij / 2                                         otherwise (a cluster: orthonormalised only),
//   e_ii = (1 - g_ii) / 2,  delta = 2 (||S - diag(lambda)||_F + max|lambda| ||R||_F),
// and writes I + E (so that U' = U (I + E)) and lambda into w (T. Ogita, K. Aishima, Japan J. Indust.
// Appl. Math. 35 (2018) 1007-1035, Algorithm 1). One workgroup per matrix; job.A = S, job.V = G,
// job.Z = I + E, job.w = lambda.
// Options: ident: U = I (S = A in J.A, G = I, not read), so that I + E is the refined U itself;
// cls (with wbase): occupation classes of the orbitals, orbital i of job J at cls[J.w - wbase + i],
// 0 occupied, 1 empty (no fractional ones): a pair within a class is only orthonormalised, a pair
// across is corrected if its first-order mixing is below ovmix (else emax = 1: not converged).
__global__ __launch_bounds__(256) void k_oa_step(const EigJob* jobs, double* emax, double* delta_keep, const int* slot,
                                                 int first, int ident, const int* cls, const double* wbase,
                                                 double ovmix) {
  __shared__ double red[16];
  __shared__ double lam[200];  // batched eigenproblems have n <= 200
  __shared__ signed char cl[200];
  const int jq = blockIdx.x;
  const EigJob J = jobs[jq];
  if (job_skipped(J.run)) return;
  const int n = J.n, tid = threadIdx.x, nt = blockDim.x;
  const double* S = J.A;
  const double* G = J.V;
  auto gel = [&](int e, bool diag) { return ident ? (diag ? 1.0 : 0.0) : G[e]; };
  for (int i = tid; i < n; i += nt) {
    lam[i] = S[i + (long long)i * n] / gel(i + i * n, true);
    J.w[i] = lam[i];
    cl[i] = cls ? (signed char)cls[(J.w - wbase) + i] : (signed char)-1;
  }
  __syncthreads();
  // element e = i + j n of this thread, advanced without division
  const int di = nt % n, dj = nt / n;
  double so = 0.0, ro = 0.0, am = 0.0;
  double delta = 0.0;  // the cluster threshold (not used with classes)
  if (!cls) {
    int i = tid % n, j = tid / n;
    for (int e = tid; e < n * n; e += nt) {
      const double sd = S[e] - (i == j ? lam[i] : 0.0);
      const double rd = (i == j ? 1.0 : 0.0) - gel(e, i == j);
      so += sd * sd;
      ro += rd * rd;
      i += di;
      j += dj;
      if (i >= n) {
        i -= n;
        j++;
      }
    }
  }
  if (!cls) {
    for (int i = tid; i < n; i += nt) am = fmax(am, fabs(lam[i]));
    so = block_sum(so, red);
    ro = block_sum(ro, red);
    // max |lambda| over the workgroup
    for (int o = 32; o > 0; o >>= 1) am = fmax(am, __shfl_xor(am, o, 64));
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = am;
    __syncthreads();
    am = 0.0;
    for (int w = 0; w < (int)(nt >> 6); w++) am = fmax(am, red[w]);
    // the first step's cluster threshold is kept for the later steps: a pair left unrotated as a
    // cluster is not rotated later from a smaller delta; pairs whose first-order mixing would
    // exceed 0.05 are treated as clusters too
    delta = 2.0 * (sqrt(so) + am * sqrt(ro));
    const int sl = slot ? slot[jq] : jq;
    if (first) {
      if (tid == 0) delta_keep[sl] = delta;
    } else {
      delta = fmax(delta, delta_keep[sl]);
    }
  }
  // emax: the largest correction (pairs left unrotated as clusters are not counted: a coupling
  // there changes neither P nor, beyond its own size, W)
  double emx = 0.0;
  {
    int i = tid % n, j = tid / n;
    for (int e = tid; e < n * n; e += nt) {
      const double g = gel(e, i == j);
      double v;
      if (i == j) {
        v = 0.5 * (1.0 - g);
      } else {
        const double gap = lam[j] - lam[i];
        const double num = S[e] - lam[j] * g;
        if (cl[i] >= 0) {
          if (cl[i] == cl[j]) {
            v = -0.5 * g;
          } else if (fabs(num) < ovmix * fabs(gap)) {
            v = num / gap;
          } else {
            v = -0.5 * g;
            emx = 1.0;
          }
        } else {
          v = (fabs(gap) > delta && fabs(num) < 0.05 * fabs(gap)) ? num / gap : -0.5 * g;
        }
      }
      emx = fmax(emx, fabs(v));
      J.Z[e] = (i == j ? 1.0 : 0.0) + v;
      i += di;
      j += dj;
      if (i >= n) {
        i -= n;
        j++;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) emx = fmax(emx, __shfl_xor(emx, o, 64));
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = emx;
  __syncthreads();
  if (tid == 0) {
    double m = 0.0;
    for (int w = 0; w < (int)(nt >> 6); w++) m = fmax(m, red[w]);
    emax[jq] = m;
  }
}

// Device control of the batched refinement (no host reads between its rounds). Flags are per
// molecule: run[m] while molecule m refines, in_t[m] = 1 when its refined vectors ended in the
// second buffer (T) instead of Z, fail[m] = 1 when it did not converge in its rounds.
// End of a round: a running job whose largest correction is within `accept` stops (its vectors are
// in the buffer of `parity`); after the last round the others stop as failures. cnt (nullable)
// accumulates [jobs refined, steps, failures].
__global__ void k_refine_round(int nj, const int* mol, int* run, const double* emax, double accept, int parity,
                               int last, int nst, int first_round, int* in_t, int* fail, unsigned long long* cnt) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nj) return;
  const int m = mol[q];
  if (!run[m]) return;
  if (cnt) {
    if (first_round) atomicAdd(&cnt[0], 1ull);
    atomicAdd(&cnt[1], (unsigned long long)nst);
  }
  if (emax[q] <= accept) {
    run[m] = 0;
    in_t[m] = parity;
  } else if (last) {
    run[m] = 0;
    fail[m] = 1;
    if (cnt) atomicAdd(&cnt[2], 1ull);
  }
}

// run[m] = src[m] for the molecules of the job list (the next refinement's participants)
// Statistics of a Jacobi launch over the solves that ran: acc[0] += sweeps, acc[1] += solves,
// acc[2] += unconverged, *work += their algorithmic flops (4/3 n^3 for the reduction and 2 n^3 for
// the vectors of each n x n matrix).
__global__ void k_sweep_acc(const int* stat, const EigJob* jobs, int n, double* acc, double* work) {
  __shared__ double red[16];
  double a = 0.0, c = 0.0, nj = 0.0, wk = 0.0;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    if (stat[q] == JACOBI_SKIPPED) continue;
    a += abs(stat[q]);
    c += stat[q] < 0;
    nj += 1.0;
    const double m = jobs[q].n;
    wk += (4.0 / 3.0 + 2.0) * m * m * m;
  }
  a = block_sum(a, red);
  c = block_sum(c, red);
  nj = block_sum(nj, red);
  wk = block_sum(wk, red);
  if (threadIdx.x == 0) {
    acc[0] += a;
    acc[1] += nj;
    acc[2] += c;
    *work += wk;
  }
}

__global__ void k_set_run(int nj, const int* mol, const int* src, int* run) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nj) run[mol[q]] = src[mol[q]];
}

// The refined vectors of the jobs whose result is in T (in_t and not failed) copied to Z, so that
// the products after the eigenproblem read Z for every molecule; pfp[m] (nullable) records whether
// molecule m ended its iteration with a converged split refinement (split[m] and not failed).
// With `active` (nullable) the molecules that are not active are left alone.
__global__ void k_refine_gather(const EigJob* jobs, const int* mol, long long t_minus_z, const int* in_t,
                                const int* fail, const int* split, int* pfp, const int* active) {
  const EigJob J = jobs[blockIdx.x];
  const int m = mol[blockIdx.x];
  if (active && !active[m]) return;
  if (pfp && threadIdx.x == 0) pfp[m] = (split[m] && !fail[m]) ? 1 : 0;
  if (!in_t[m] || fail[m]) return;
  const long long nn = (long long)J.n * J.n;
  const double* src = J.Z + t_minus_z;
  for (long long e = threadIdx.x; e < nn; e += blockDim.x) J.Z[e] = src[e];
}

// The refinement flags of an SCC iteration (flags = [split | full | fail | in_t], nm each) for the
// molecules mol[0..nj): split = active and allowed to refine for the occupied / empty split
// (ok[m], when use_split); full = active, not split, FP32 path; fail = active, not split, FP64 path
// (the FP64 Jacobi); in_t = 0.
__global__ void k_iter_flags(int nj, const int* mol, const int* active, const int* ok, int use_split, int fp32, int nm,
                             int* flags) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nj) return;
  const int m = mol[q];
  const int a = active[m] != 0, sp = a && use_split && ok[m] != 0;
  flags[m] = sp;
  flags[nm + m] = (a && !sp && fp32) ? 1 : 0;
  flags[2 * nm + m] = (a && !sp && !fp32) ? 1 : 0;
  flags[3 * nm + m] = 0;
}

// Whether molecule m may refine its previous orbitals for the occupied / empty split in the next
// SCC iteration: no fractional occupation and the last residual (mst[8 m + 2]) below `gate`;
// ok[m] = 0 for molecules not in mlist.
__global__ void k_split_ok(const int* mlist, int nm, const int* nfrac, const double* mst, double gate, int* ok) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nm) return;
  const int m = mlist[q];
  ok[m] = (nfrac[m] == 0 && mst[8 * m + 2] < gate) ? 1 : 0;
}

// The Ogita-Aishima step of k_oa_step for one large matrix, over many workgroups:
// k_oa_big_norms: lambda_i = s_ii / g_ii (workgroup 0 writes them) and per-workgroup partial sums
// of ||S - diag(lambda)||_F^2, ||I - G||_F^2 and max |lambda| (part[3 * block + 0..2]);
// k_oa_big_e: delta from the partials (every workgroup the same, fixed order), then I + E
// elementwise; emax_part[block] = max |e_ij| of the workgroup.
__global__ __launch_bounds__(256) void k_oa_big_norms(const EigJob J, double* part) {
  __shared__ double red[16];
  const int n = J.n;
  const long long nn = (long long)n * n;
  double so = 0.0, ro = 0.0, am = 0.0;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    const double s = J.A[e], g = J.V[e];
    double sd = s, rd = -g;
    if (i == j) {
      const double l = s / g;
      sd = s - l;
      rd = 1.0 - g;
      am = fmax(am, fabs(l));
      J.w[i] = l;  // every workgroup writes its own diagonal entries
    }
    so += sd * sd;
    ro += rd * rd;
  }
  so = block_sum(so, red);
  ro = block_sum(ro, red);
  for (int o = 32; o > 0; o >>= 1) am = fmax(am, __shfl_xor(am, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = am;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) m = fmax(m, red[w]);
    part[3 * blockIdx.x] = so;
    part[3 * blockIdx.x + 1] = ro;
    part[3 * blockIdx.x + 2] = m;
  }
}

// cls (nullable): per column, 0 = fully occupied, 1 = empty, 2 = fractional occupation. With it,
// an occupied-empty pair is always rotated (its gap is the occupied-virtual gap; a first-order
// mixing >= ovmix marks the step as not convergent, emax >= 1), pairs inside the occupied or the
// empty block are only orthogonalised (rotations there leave P = C f C^T unchanged when f is
// constant on the block), and pairs with a fractional column follow the delta rule.
__global__ __launch_bounds__(256) void k_oa_big_e(const EigJob J, const double* part, int nparts, double* delta_keep,
                                                  int first, double* emax_part, const int* cls, double ovmix) {
  __shared__ double red[16];
  __shared__ double s_delta;
  const int n = J.n;
  const long long nn = (long long)n * n;
  if (threadIdx.x == 0) {
    double so = 0.0, ro = 0.0, am = 0.0;
    for (int q = 0; q < nparts; q++) {
      so += part[3 * q];
      ro += part[3 * q + 1];
      am = fmax(am, part[3 * q + 2]);
    }
    double d = 2.0 * (sqrt(so) + am * sqrt(ro));
    if (!first) d = fmax(d, *delta_keep);
    s_delta = d;
  }
  __syncthreads();
  const double delta = s_delta;
  if (first && blockIdx.x == 0 && threadIdx.x == 0) *delta_keep = delta;
  double emx = 0.0;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    const double g = J.V[e];
    double v;
    if (i == j) {
      v = 0.5 * (1.0 - g);
    } else {
      const double gap = J.w[j] - J.w[i];
      const double num = J.A[e] - J.w[j] * g;
      const int ci = cls ? cls[i] : 2, cj = cls ? cls[j] : 2;
      if (ci != 2 && cj != 2) {
        if (ci == cj) {
          v = -0.5 * g;
        } else if (fabs(num) < ovmix * fabs(gap)) {
          v = num / gap;
        } else {
          v = -0.5 * g;
          emx = fmax(emx, 1.0);
        }
      } else if (fabs(gap) > delta && fabs(num) < 0.05 * fabs(gap)) {
        v = num / gap;
      } else {
        v = -0.5 * g;
        emx = fmax(emx, fabs(num));
      }
    }
    emx = fmax(emx, fabs(v));
    J.Z[e] = (i == j ? 1.0 : 0.0) + v;
  }
  for (int o = 32; o > 0; o >>= 1) emx = fmax(emx, __shfl_xor(emx, o, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = emx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) m = fmax(m, red[w]);
    emax_part[blockIdx.x] = m;
  }
}

// Square sub-blocks copied out of column-major matrices, one workgroup per block:
// dst[r + c nb] = src[r + c ld] for r, c < nb.
struct BlockCopy {
  const double* src;
  double* dst;
  int nb, ld;
};
__global__ void k_copy_blocks(const BlockCopy* jobs) {
  const BlockCopy J = jobs[blockIdx.x];
  const int nn = J.nb * J.nb;
  for (int e = threadIdx.x; e < nn; e += blockDim.x) J.dst[e] = J.src[(e % J.nb) + (long long)(e / J.nb) * J.ld];
}

// Occupation class of each orbital column for k_oa_big_e: 0 if f > 2 - 1e-12, 1 if f < 1e-12,
// else 2; *nfrac counts the class-2 columns.
__global__ void k_occ_class(const double* f, int n, int* cls, int* nfrac) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = f[i] > 2.0 - 1e-12 ? 0 : (f[i] < 1e-12 ? 1 : 2);
  cls[i] = c;
  if (c == 2) atomicAdd(nfrac, 1);
}

// S-orthonormal start basis B0 = L^-T from the Cholesky factor S = L L^T (the reduction of
// GenEigSolver, base/src/linalg/eigen.rs:72-113), one workgroup per matrix in LDS (n^2 + n doubles):
// L in the lower triangle, then X = L^-1 by right-looking forward substitution with its strictly
// lower part stored transposed in the upper triangle (X(i, j) at [j + i n], which is where B0 = X^T
// keeps it) and its diagonal in xd. job.A = S, job.Z = B0 (upper triangular); fail[job] = 1 if S
// is not positive definite.
__global__ __launch_bounds__(512) void k_chol_basis(const EigJob* jobs, int* fail) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ int bad;
  const EigJob J = jobs[blockIdx.x];
  const int n = J.n, tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
  double* M = (double*)lds_raw;
  double* xd = M + n * n;
  for (int e = tid; e < n * n; e += nt) M[e] = J.A[e];
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int k = 0; k < n; k++) {
    if (tid == 0) {
      double d = M[k + k * n];
      if (!(d > 0.0)) {
        bad = 1;
        d = 1.0;
      }
      M[k + k * n] = sqrt(d);
    }
    __syncthreads();
    const double lkk = M[k + k * n];
    for (int i = k + 1 + tid; i < n; i += nt) M[i + k * n] /= lkk;
    __syncthreads();
    // trailing lower triangle: column j by one wavefront, rows i >= j over the lanes
    for (int j = k + 1 + wv; j < n; j += nw) {
      const double ljk = M[j + k * n];
      for (int i = j + lane; i < n; i += 64) M[i + j * n] -= M[i + k * n] * ljk;
    }
    __syncthreads();
  }
  // strictly upper part <- 0 (X^T lives there), xd <- 1
  for (int e = tid; e < n * n; e += nt)
    if (e % n < e / n) M[e] = 0.0;
  for (int i = tid; i < n; i += nt) xd[i] = 1.0;
  __syncthreads();
  for (int k = 0; k < n; k++) {
    const double lkk = M[k + k * n];
    // row k of X is final once divided by L(k, k)
    for (int j = tid; j <= k; j += nt) {
      if (j == k)
        xd[k] /= lkk;
      else
        M[j + k * n] /= lkk;
    }
    __syncthreads();
    // X(i, j) -= L(i, k) X(k, j) for i > k, j <= k: row i by one wavefront, columns over the lanes
    for (int i = k + 1 + wv; i < n; i += nw) {
      const double lik = M[i + k * n];
      for (int j = lane; j <= k; j += 64) M[j + i * n] -= lik * (j == k ? xd[k] : M[j + k * n]);
    }
    __syncthreads();
  }
  for (int e = tid; e < n * n; e += nt) {
    const int i = e % n, j = e / n;
    J.Z[e] = i < j ? M[e] : (i == j ? xd[i] : 0.0);
  }
  if (tid == 0) fail[blockIdx.x] = bad;
}

// The same for matrices whose full square does not fit in LDS (n <= 200): the lower triangle packed
// by columns (element (r, c), r >= c, at c (2n - c + 1) / 2 + r - c) plus one column, n (n + 1) / 2 + n
// doubles. L by the right-looking Cholesky, then X = L^-1 in place, columns from the last:
// X(j, j) = 1 / L(j, j), X(i, j) = -X(j, j) sum_{k = j+1..i} X(i, k) L(k, j) for i > j (column j of L
// copied aside first; the columns k > j of X are final). job.Z = B0 = X^T (upper triangular).
__global__ __launch_bounds__(512) void k_chol_basis_packed(const EigJob* jobs, int* fail) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ int bad;
  const EigJob J = jobs[blockIdx.x];
  const int n = J.n, tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
  double* M = (double*)lds_raw;
  double* col = M + n * (n + 1) / 2;
  auto pidx = [&](int r, int c) { return c * (2 * n - c + 1) / 2 + (r - c); };
  for (int c = wv; c < n; c += nw)
    for (int r = c + lane; r < n; r += 64) M[pidx(r, c)] = J.A[r + (long long)c * n];
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int k = 0; k < n; k++) {
    const int pkk = pidx(k, k);
    if (tid == 0) {
      double d = M[pkk];
      if (!(d > 0.0)) {
        bad = 1;
        d = 1.0;
      }
      M[pkk] = sqrt(d);
    }
    __syncthreads();
    const double lkk = M[pkk];
    for (int i = k + 1 + tid; i < n; i += nt) M[pkk + (i - k)] /= lkk;
    __syncthreads();
    // trailing lower triangle: column j by one wavefront, rows i >= j over the lanes
    for (int j = k + 1 + wv; j < n; j += nw) {
      const int pj = pidx(j, j), pk = pkk + (j - k);  // (j, j) and (j, k): rows i >= j follow
      const double ljk = M[pk];
      for (int i = j + lane; i < n; i += 64) M[pj + (i - j)] -= M[pk + (i - j)] * ljk;
    }
    __syncthreads();
  }
  for (int j = n - 1; j >= 0; j--) {
    const int pjj = pidx(j, j);
    for (int i = j + 1 + tid; i < n; i += nt) col[i] = M[pjj + (i - j)];
    __syncthreads();
    const double xjj = 1.0 / M[pjj];
    for (int i = j + 1 + tid; i < n; i += nt) {
      double s = 0.0;
      int p = pidx(i, j + 1);
      for (int k = j + 1; k <= i; k++) {
        s += M[p] * col[k];
        p += n - k - 1;
      }
      M[pjj + (i - j)] = -xjj * s;
    }
    __syncthreads();
    if (tid == 0) M[pjj] = xjj;
    __syncthreads();
  }
  for (int e = tid; e < n * n; e += nt) {
    const int r = e % n, c = e / n;
    J.Z[e] = c >= r ? M[pidx(c, r)] : 0.0;
  }
  if (tid == 0) fail[blockIdx.x] = bad;
}

// EEQ charges of the listed molecules by LU with partial pivoting (the first maximum, as the
// reference's dense LU in base/src/eeq.rs), in LDS ((n+1)^2 + n+1 doubles), then the start shell
// charges as in k_eeq_solve; a zero pivot or a non-finite charge gives the neutral start.
__global__ __launch_bounds__(512) void k_eeq_lu(Batch b, const int* mlist, const long long* eq0, const double* Ag,
                                                const double* rhs, double* qstart, int* eeq_failed) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ int bad, piv;
  __shared__ double pv[8];
  __shared__ int pi[8];
  const int m = mlist[blockIdx.x];
  const Mol M = b.mol[m];
  const int n = M.nat + 1, tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
  double* A = (double*)lds_raw;
  double* x = A + n * n;
  for (int e = tid; e < n * n; e += nt) A[e] = Ag[eq0[m] + e];
  for (int i = tid; i < n; i += nt) x[i] = rhs[(long long)M.at0 + m + i];
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int k = 0; k < n; k++) {
    // pivot: the first row i >= k of maximum |A(i, k)|
    double best = -1.0;
    int bi = n;
    for (int i = k + tid; i < n; i += nt) {
      const double v = fabs(A[i + k * n]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ov > best || (ov == best && oi < bi)) {
        best = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      pv[wv] = best;
      pi[wv] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      double bv = pv[0];
      int b0 = pi[0];
      for (int w = 1; w < nw; w++)
        if (pv[w] > bv || (pv[w] == bv && pi[w] < b0)) {
          bv = pv[w];
          b0 = pi[w];
        }
      piv = b0;
      if (!(bv > 0.0)) bad = 1;
    }
    __syncthreads();
    if (bad) break;
    const int p = piv;
    if (p != k) {
      for (int j = tid; j < n; j += nt) {
        const double t = A[k + j * n];
        A[k + j * n] = A[p + j * n];
        A[p + j * n] = t;
      }
      if (tid == 0) {
        const double t = x[k];
        x[k] = x[p];
        x[p] = t;
      }
      __syncthreads();
    }
    const double akk = A[k + k * n];
    for (int i = k + 1 + tid; i < n; i += nt) A[i + k * n] /= akk;
    __syncthreads();
    // trailing update: column j by one wavefront, rows over the lanes
    for (int j = k + 1 + wv; j < n; j += nw) {
      const double akj = A[k + j * n];
      for (int i = k + 1 + lane; i < n; i += 64) A[i + j * n] -= A[i + k * n] * akj;
    }
    __syncthreads();
  }
  if (!bad) {
    // L y = P b (unit lower), then U x = y
    for (int k = 0; k < n; k++) {
      const double xk = x[k];
      for (int i = k + 1 + tid; i < n; i += nt) x[i] -= A[i + k * n] * xk;
      __syncthreads();
    }
    for (int k = n - 1; k >= 0; k--) {
      if (tid == 0) x[k] /= A[k + k * n];
      __syncthreads();
      const double xk = x[k];
      for (int i = tid; i < k; i += nt) x[i] -= A[i + k * n] * xk;
      __syncthreads();
    }
  }
  for (int i = tid; i < M.nat; i += nt)
    if (!isfinite(x[i])) bad = 1;
  __syncthreads();
  for (int i = tid; i < M.nat; i += nt) {
    const int a = M.at0 + i;
    double tot = 0.0;
    for (int s = b.a.ish[a]; s < b.a.ish[a] + b.a.nsh[a]; s++) tot += b.s.refocc[s];
    for (int s = b.a.ish[a]; s < b.a.ish[a] + b.a.nsh[a]; s++)
      qstart[s] = (bad || !(tot > 0.0)) ? 0.0 : x[i] * b.s.refocc[s] / tot;
  }
  if (tid == 0 && bad) eeq_failed[m] = 1;
}

// dynamic LDS bytes of k_jacobi<mode, T, pack> for an n x n matrix in elements of `es` bytes
inline size_t jacobi_lds_p(int n, int mode, size_t es, bool pack) {
  const size_t ld = pack ? n : (n | 1), ldv = (mode == 2 && !pack) ? ld : n;
  const size_t np = (n + 1) / 2, na = pack ? (size_t)n * (n + 1) / 2 : (size_t)n * ld;
  const size_t mats = (size_t)(mode >= 1 ? 1 : 0) * na * es + (size_t)(mode == 2 ? 1 : 0) * n * ldv * es;
  return (mats + 15) / 16 * 16 + 2 * np * (es == 4 ? 16 : 32);  // + two buffers of JRot records
}
// the layout of the instantiations: A packed except in the FP32 LDS mode 2 (k_jacobi2_f32, where the
// packed index arithmetic costs more than two workgroups per CU gain for n = 100-115, measured)
inline size_t jacobi_lds(int n, int mode, size_t es = 8) { return jacobi_lds_p(n, mode, es, !(mode == 2 && es == 4)); }

}  // namespace axtb
