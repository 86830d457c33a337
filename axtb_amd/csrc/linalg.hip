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
};

typedef double d4v __attribute__((ext_vector_type(4)));

constexpr int GT = 64;   // tile edge
constexpr int GK = 16;   // k step

// 256 threads = 4 wavefronts; wavefront w owns the 32 x 32 quadrant (w & 1, w >> 1) of the tile
// as 2 x 2 MFMA blocks of 16 x 16. LDS holds op(A) as As[k][row] and op(B) as Bs[k][col].
__global__ __launch_bounds__(256) void k_gemm(const GemmJob* jobs) {
  const GemmJob J = jobs[blockIdx.z];
  const int row0 = blockIdx.y * GT, col0 = blockIdx.x * GT;
  if (row0 >= J.m || col0 >= J.n) return;
  __shared__ double As[GK][GT + 1];
  __shared__ double Bs[GK][GT + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = (w & 1) * 32, wc = (w >> 1) * 32;
  d4v acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = (d4v){0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < J.k; k0 += GK) {
    // stage op(A)[row0:+64, k0:+16] and op(B)[k0:+16, col0:+64] (zero padded)
    for (int e = tid; e < GT * GK; e += 256) {
      int r, kk;
      if (J.ta == 0) {
        r = e % GT;
        kk = e / GT;
      } else {
        kk = e % GK;
        r = e / GK;
      }
      int gr = row0 + r, gk = k0 + kk;
      double v = 0.0;
      if (gr < J.m && gk < J.k) {
        v = J.ta == 0 ? J.A[gr + gk * J.lda] : J.A[gk + (long long)gr * J.lda];
        if (J.dk) v *= J.dk[gk];
      }
      As[kk][r] = v;
    }
    for (int e = tid; e < GT * GK; e += 256) {
      int c, kk;
      if (J.tb == 0) {
        kk = e % GK;
        c = e / GK;
      } else {
        c = e % GT;
        kk = e / GT;
      }
      int gc = col0 + c, gk = k0 + kk;
      double v = 0.0;
      if (gc < J.n && gk < J.k) v = J.tb == 0 ? J.B[gk + (long long)gc * J.ldb] : J.B[gc + (long long)gk * J.ldb];
      Bs[kk][c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      const int kl = kk + (lane >> 4), li = lane & 15;
      double a0 = As[kl][wr + li], a1 = As[kl][wr + 16 + li];
      double b0 = Bs[kl][wc + li], b1 = Bs[kl][wc + 16 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
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

// ------------------------------------------------------------------ Jacobi eigensolver
struct EigJob {
  int n;
  double* A;      // n x n (ld n), destroyed
  double* V;      // n x n work (ld n)
  double* w;      // eigenvalues, ascending
  double* Z;      // eigenvectors (ld n), columns in the order of w
};

// Pair k of round r of the round-robin tournament over nn (even) indices: index 0 stays, the
// others rotate.
__device__ __forceinline__ void tour_pair(int r, int k, int nn, int& p, int& q) {
  auto pos = [&](int s) { return s == 0 ? 0 : 1 + (s - 1 + r) % (nn - 1); };
  p = pos(k);
  q = pos(nn - 1 - k);
  if (p > q) {
    int t = p;
    p = q;
    q = t;
  }
}

// One workgroup per matrix. USE_LDS: the matrix is copied to LDS (n*n doubles must fit).
// status[job] = number of sweeps (negative if not converged within max_sweeps).
template <bool USE_LDS>
__global__ __launch_bounds__(256) void k_jacobi(const EigJob* jobs, int max_sweeps, int* status) {
  extern __shared__ double lds[];
  __shared__ int rotated;
  __shared__ double anorm;
  const EigJob J = jobs[blockIdx.x];
  const int n = J.n, tid = threadIdx.x, nt = blockDim.x;
  if (n <= 0) return;
  // dynamic LDS: [A (n*n, USE_LDS only)] [c, s of two rounds: 4*npmax] [p, q of two rounds]
  const int npmax = (n + 1) / 2;
  double* A = USE_LDS ? lds : J.A;
  double* csb = lds + (USE_LDS ? n * n : 0);
  int* pqb = (int*)(csb + 4 * npmax);
#define CS(b, x, k) csb[((b)*2 + (x)) * npmax + (k)]
#define PQ(b, x, k) pqb[((b)*2 + (x)) * npmax + (k)]
  double* V = J.V;
  if (USE_LDS)
    for (int e = tid; e < n * n; e += nt) A[e] = J.A[e];
  for (int e = tid; e < n * n; e += nt) V[e] = (e % n == e / n) ? 1.0 : 0.0;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s = fmax(s, fabs(A[i + i * n]));
    for (int e = 0; e < n * n; e++) s = fmax(s, fabs(A[e]));
    anorm = s;
  }
  const int nn = n + (n & 1), np = nn / 2;
  int sweep = 0;
  bool conv = n == 1;
  for (sweep = 0; sweep < max_sweeps && !conv; sweep++) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int r = 0; r < nn - 1; r++) {
      const int bsel = r & 1;
      for (int k = tid; k < np; k += nt) {
        int p, q;
        tour_pair(r, k, nn, p, q);
        double c = 1.0, s = 0.0;
        int ok = 0;
        if (q < n) {
          double app = A[p + p * n], aqq = A[q + q * n], apq = A[p + q * n];
          double t1 = 1e-15 * sqrt(fabs(app) * fabs(aqq)), t2 = 1e-17 * anorm;
          if (fabs(apq) > t1 && fabs(apq) > t2) {
            double tau = (aqq - app) / (2.0 * apq);
            double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
            ok = 1;
          }
        }
        CS(bsel, 0, k) = c;
        CS(bsel, 1, k) = s;
        PQ(bsel, 0, k) = p;
        PQ(bsel, 1, k) = ok ? q : -1;
        if (ok) rotated = 1;
      }
      __syncthreads();
      // A <- A J (columns p, q of every rotated pair)
      for (int e = tid; e < np * n; e += nt) {
        int k = e / n, i = e % n;
        int q = PQ(bsel, 1, k);
        if (q < 0) continue;
        int p = PQ(bsel, 0, k);
        double c = CS(bsel, 0, k), s = CS(bsel, 1, k);
        double aip = A[i + p * n], aiq = A[i + q * n];
        A[i + p * n] = c * aip - s * aiq;
        A[i + q * n] = s * aip + c * aiq;
      }
      __syncthreads();
      // A <- J^T A (rows p, q)
      for (int e = tid; e < np * n; e += nt) {
        int k = e / n, j = e % n;
        int q = PQ(bsel, 1, k);
        if (q < 0) continue;
        int p = PQ(bsel, 0, k);
        double c = CS(bsel, 0, k), s = CS(bsel, 1, k);
        double apj = A[p + j * n], aqj = A[q + j * n];
        A[p + j * n] = c * apj - s * aqj;
        A[q + j * n] = s * apj + c * aqj;
      }
      __syncthreads();
      for (int k = tid; k < np; k += nt) {
        int q = PQ(bsel, 1, k);
        if (q < 0) continue;
        int p = PQ(bsel, 0, k);
        A[p + q * n] = 0.0;
        A[q + p * n] = 0.0;
      }
      // V <- V J (reads this round's rotations; the next round writes the other buffer)
      for (int e = tid; e < np * n; e += nt) {
        int k = e / n, i = e % n;
        int q = PQ(bsel, 1, k);
        if (q < 0) continue;
        int p = PQ(bsel, 0, k);
        double c = CS(bsel, 0, k), s = CS(bsel, 1, k);
        double vip = V[i + p * n], viq = V[i + q * n];
        V[i + p * n] = c * vip - s * viq;
        V[i + q * n] = s * vip + c * viq;
      }
      __syncthreads();
    }
    conv = rotated == 0;
    __syncthreads();
  }
  // eigenvalues ascending (ties by index), vectors permuted alike
  for (int i = tid; i < n; i += nt) {
    double di = A[i + i * n];
    int rank = 0;
    for (int j = 0; j < n; j++) {
      double dj = A[j + j * n];
      if (dj < di || (dj == di && j < i)) rank++;
    }
    J.w[rank] = di;
    for (int r = 0; r < n; r++) J.Z[r + (long long)rank * n] = V[r + (long long)i * n];
  }
  if (tid == 0) status[blockIdx.x] = conv ? sweep : -sweep;
#undef CS
#undef PQ
}

// dynamic LDS bytes of k_jacobi for an n x n matrix
inline size_t jacobi_lds(int n, bool use_lds) {
  size_t np = (n + 1) / 2;
  return (use_lds ? (size_t)n * n * 8 : 0) + np * 4 * 8 + np * 4 * 4;
}

template __global__ void k_jacobi<true>(const EigJob*, int, int*);
template __global__ void k_jacobi<false>(const EigJob*, int, int*);

}  // namespace axtb
