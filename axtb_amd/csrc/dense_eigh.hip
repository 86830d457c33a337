// Dense symmetric eigensolver for large matrices (FP64): Householder reduction to tridiagonal
// form, divide and conquer on the tridiagonal matrix (Cuppen's tearing, Gu-Eisenstat eigenvectors
// with deflation), and the back-transformation of the eigenvectors with the reflectors in blocked
// (WY) form on the matrix cores. All matrices column-major.
#include "common.hpp"

namespace axtb {

// ------------------------------------------------------------------ Householder tridiagonalisation
// Reflector of column i: x = A[i+1:n, i]; H = I - tau v v^T with v[0] = 1 maps x to (beta, 0, ...).
// v is stored in A[i+1:n, i], d[i] = A[i][i], e[i] = beta. One workgroup.
__global__ void k_hh_reflector(double* A, int n, int i, double* tau, double* d, double* e) {
  __shared__ double red[16];
  const int m = n - i - 1;
  double* x = A + (long long)i * n + i + 1;
  double s = 0.0;
  for (int k = 1 + threadIdx.x; k < m; k += blockDim.x) s += x[k] * x[k];
  s = block_sum(s, red);
  const double alpha = x[0];
  if (threadIdx.x == 0) d[i] = A[(long long)i * n + i];
  if (s == 0.0) {
    // nothing to annihilate: H = I
    if (threadIdx.x == 0) {
      tau[i] = 0.0;
      e[i] = alpha;
      x[0] = 1.0;
    }
    return;
  }
  const double nrm = sqrt(alpha * alpha + s);
  const double beta = alpha >= 0.0 ? -nrm : nrm;
  const double scal = 1.0 / (alpha - beta);
  for (int k = 1 + threadIdx.x; k < m; k += blockDim.x) x[k] *= scal;
  if (threadIdx.x == 0) {
    tau[i] = (beta - alpha) / beta;
    e[i] = beta;
    x[0] = 1.0;
  }
}

// Partial products of p = A22 v (A22 = A[i+1:, i+1:]) over column chunks of 64: part[chunk][row].
__global__ void k_symv_part(const double* A, int n, int i, double* part) {
  const int m = n - i - 1;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c0 = blockIdx.y * 64;
  __shared__ double vs[64];
  const double* v = A + (long long)i * n + i + 1;
  if (threadIdx.x < 64) vs[threadIdx.x] = c0 + (int)threadIdx.x < m ? v[c0 + threadIdx.x] : 0.0;
  __syncthreads();
  if (r >= m) return;
  const double* a = A + (long long)(i + 1) * n + (i + 1);
  double s = 0.0;
  const int c1 = min(c0 + 64, m);
  for (int c = c0; c < c1; c++) s += a[(long long)c * n + r] * vs[c - c0];
  part[(long long)blockIdx.y * m + r] = s;
}

// p = tau * sum of the chunks, then w = p - (tau/2)(p.v) v (one workgroup).
__global__ void k_hh_w(const double* A, int n, int i, const double* tau, const double* part, int nchunk, double* w) {
  __shared__ double red[16];
  const int m = n - i - 1;
  const double t = tau[i];
  const double* v = A + (long long)i * n + i + 1;
  double dot = 0.0;
  for (int r = threadIdx.x; r < m; r += blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < nchunk; c++) s += part[(long long)c * m + r];
    s *= t;
    w[r] = s;
    dot += s * v[r];
  }
  dot = block_sum(dot, red);
  const double f = 0.5 * t * dot;
  for (int r = threadIdx.x; r < m; r += blockDim.x) w[r] -= f * v[r];
}

// A22 -= v w^T + w v^T
__global__ void k_syr2(double* A, int n, int i, const double* w) {
  const int m = n - i - 1;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= m || c >= m) return;
  const double* v = A + (long long)i * n + i + 1;
  double* a = A + (long long)(i + 1 + c) * n + (i + 1);
  a[r] -= v[r] * w[c] + w[r] * v[c];
}

__global__ void k_hh_last(const double* A, int n, double* d, double* e, double* tau) {
  // the last diagonal element; no reflector for the last column
  d[n - 1] = A[(long long)(n - 1) * n + (n - 1)];
  if (n >= 2) {
    // column n-2 has a 1-element x: its reflector was formed by k_hh_reflector
  }
  tau[n - 1] = 0.0;
  e[n - 1] = 0.0;
}

// ------------------------------------------------------------------ divide and conquer
struct DcNode {
  int s, m, m1;       // rows [s, s+m), split after m1 rows
};

// Dense leaf matrices (tridiagonal, torn) for the Jacobi solver: leaf l of size m at s.
__global__ void k_dc_leaves(const int* ls, const int* lm, const double* dt, const double* e, double* L, int lmax) {
  const int l = blockIdx.x;
  const int s = ls[l], m = lm[l];
  double* out = L + (long long)l * lmax * lmax;
  for (int x = threadIdx.x; x < m * m; x += blockDim.x) {
    int i = x % m, j = x / m;
    double v = 0.0;
    if (i == j) v = dt[s + i];
    else if (i == j + 1) v = e[s + j];
    else if (j == i + 1) v = e[s + i];
    out[i + j * m] = v;
  }
}

// Scatter the leaf eigenpairs into Q (block diagonal) and D.
__global__ void k_dc_scatter(const int* ls, const int* lm, const double* Z, const double* w, int lmax, double* Q, int n,
                             double* D) {
  const int l = blockIdx.x;
  const int s = ls[l], m = lm[l];
  const double* z = Z + (long long)l * lmax * lmax;
  for (int x = threadIdx.x; x < m * m; x += blockDim.x) {
    int i = x % m, j = x / m;
    Q[(long long)(s + j) * n + s + i] = z[i + j * m];
  }
  for (int j = threadIdx.x; j < m; j += blockDim.x) D[s + j] = w[(long long)l * lmax + j];
}

// Merge preparation, one workgroup per node: the sorted poles d (merged from the two sorted
// children), the normalised z, rho', and deflation (sequential scan by thread 0). Outputs per node
// at offset s: ds[s+j] sorted poles, zs[s+j], perm[s+j] = column of Q of sorted entry j,
// kind[s+j] = 1 kept (non-deflated) / 0 deflated, and the Givens rotations (rot_i, rot_j, c, s)
// applied to deflate close poles. nk[node] = kept count, nr[node] = rotation count,
// rho[node] = rho'.
__global__ void k_dc_prep(const DcNode* nodes, const double* Q, int n, const double* D, const double* e, double* ds,
                          double* zs, int* perm, int* kind, int* rot_ij, double* rot_cs, int* nk, int* nr,
                          double* rho_out) {
  const DcNode N = nodes[blockIdx.x];
  const int s = N.s, m = N.m, m1 = N.m1;
  const double beta = e[s + m1 - 1];
  __shared__ double red[16];
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    perm[s + j] = s + j;
    kind[s + j] = 0;
  }
  __syncthreads();
  // sorted merge of D[s:s+m1] and D[s+m1:s+m] (each ascending): rank of every entry
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    double v = D[s + j];
    int pos;
    if (j < m1) {
      // number of right entries < v
      int lo = 0, hi = m - m1;
      while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (D[s + m1 + mid] < v) lo = mid + 1;
        else hi = mid;
      }
      pos = j + lo;
    } else {
      int jr = j - m1;
      int lo = 0, hi = m1;
      while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (D[s + mid] <= v) lo = mid + 1;
        else hi = mid;
      }
      pos = jr + lo;
    }
    // z: last row of the left block's vectors, sign(beta) * first row of the right block's
    double z = j < m1 ? Q[(long long)(s + j) * n + (s + m1 - 1)] : (beta >= 0.0 ? 1.0 : -1.0) * Q[(long long)(s + j) * n + (s + m1)];
    ds[s + pos] = v;
    zs[s + pos] = z;
    perm[s + pos] = s + j;
  }
  __syncthreads();
  double nz = 0.0;
  for (int j = threadIdx.x; j < m; j += blockDim.x) nz += zs[s + j] * zs[s + j];
  nz = block_sum(nz, red);
  const double rho = fabs(beta) * nz;
  const double inv = 1.0 / sqrt(nz);
  for (int j = threadIdx.x; j < m; j += blockDim.x) zs[s + j] *= inv;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double dmax = 0.0;
  for (int j = 0; j < m; j++) dmax = fmax(dmax, fabs(ds[s + j]));
  const double eps = 2.220446049250313e-16;
  const double tol = 8.0 * eps * fmax(dmax, rho);
  int kept = 0, nrot = 0, pj = -1;
  for (int j = 0; j < m; j++) {
    double zj = zs[s + j];
    if (rho * fabs(zj) <= tol) {
      kind[s + j] = 0;
      continue;
    }
    if (pj < 0) {
      pj = j;
      continue;
    }
    double zp = zs[s + pj];
    double t = hypot(zp, zj);
    double c = zj / t, sn = -zp / t;
    if (fabs((ds[s + j] - ds[s + pj]) * c * sn) <= tol) {
      // rotate so that z_pj = 0: pj deflated with the rotated pole, j keeps t
      double dp = ds[s + pj], dj = ds[s + j];
      ds[s + pj] = c * c * dp + sn * sn * dj;
      ds[s + j] = sn * sn * dp + c * c * dj;
      zs[s + pj] = 0.0;
      zs[s + j] = t;
      kind[s + pj] = 0;
      rot_ij[2 * (s + nrot)] = perm[s + pj];
      rot_ij[2 * (s + nrot) + 1] = perm[s + j];
      rot_cs[2 * (s + nrot)] = c;
      rot_cs[2 * (s + nrot) + 1] = sn;
      nrot++;
      pj = j;
    } else {
      kind[s + pj] = 1;
      kept++;
      pj = j;
    }
  }
  if (pj >= 0) {
    kind[s + pj] = 1;
    kept++;
  }
  nk[blockIdx.x] = kept;
  nr[blockIdx.x] = nrot;
  rho_out[blockIdx.x] = rho;
}

// Apply the deflation rotations of each node to the columns of Q (rows of the node), in order:
// [q_i, q_j] <- [c q_i + s q_j, -s q_i + c q_j].
__global__ void k_dc_rot(const DcNode* nodes, const int* nr, const int* rot_ij, const double* rot_cs, double* Q, int n) {
  const DcNode N = nodes[blockIdx.y];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= N.m) return;
  const int row = N.s + r;
  for (int k = 0; k < nr[blockIdx.y]; k++) {
    int ci = rot_ij[2 * (N.s + k)], cj = rot_ij[2 * (N.s + k) + 1];
    double c = rot_cs[2 * (N.s + k)], sn = rot_cs[2 * (N.s + k) + 1];
    double qi = Q[(long long)ci * n + row], qj = Q[(long long)cj * n + row];
    Q[(long long)ci * n + row] = c * qi + sn * qj;
    Q[(long long)cj * n + row] = -sn * qi + c * qj;
  }
}

// Compact the kept poles of each node: kd/kz/kcol at s + (kept index), in ascending order.
__global__ void k_dc_compact(const DcNode* nodes, const int* kind, const double* ds, const double* zs, const int* perm,
                             double* kd, double* kz, int* kcol) {
  const DcNode N = nodes[blockIdx.x];
  if (threadIdx.x != 0) return;
  int k = 0;
  for (int j = 0; j < N.m; j++)
    if (kind[N.s + j]) {
      kd[N.s + k] = ds[N.s + j];
      kz[N.s + k] = zs[N.s + j];
      kcol[N.s + k] = perm[N.s + j];
      k++;
    }
}

// Secular equation 1 + rho sum_i z_i^2 / (d_i - lambda) = 0 of each kept root j: lambda_j in
// (d_j, d_{j+1}) (the last in (d_K-1, d_K-1 + rho)), found by bisection in tau = lambda - d_org
// relative to the nearer pole (org), which keeps lambda - d_i accurate.
__global__ void k_dc_secular(const DcNode* nodes, const int* nk, const double* rho_, const double* kd,
                             const double* kz, double* tau, int* org) {
  const DcNode N = nodes[blockIdx.y];
  const int K = nk[blockIdx.y];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= K) return;
  const double rho = rho_[blockIdx.y];
  const double* d = kd + N.s;
  const double* z = kz + N.s;
  const double lo_d = d[j];
  const double hi_d = j + 1 < K ? d[j + 1] : d[j] + rho;
  // f at the midpoint decides the origin
  const double mid = 0.5 * (hi_d - lo_d);
  double f = 1.0 / rho;
  for (int i = 0; i < K; i++) f += z[i] * z[i] / ((d[i] - lo_d) - mid);
  int o;
  double lo, hi;
  if (f >= 0.0 || j + 1 >= K) {
    o = j;  // root in (d_j, mid]
    lo = 0.0;
    hi = j + 1 < K ? mid : rho;
  } else {
    o = j + 1;  // root in [mid, d_j+1)
    lo = -mid;
    hi = 0.0;
  }
  const double dorg = d[o];
  for (int it = 0; it < 200; it++) {
    double t = 0.5 * (lo + hi);
    if (t == lo || t == hi) break;
    double g = 1.0 / rho;
    for (int i = 0; i < K; i++) g += z[i] * z[i] / ((d[i] - dorg) - t);
    if (g > 0.0) hi = t;
    else lo = t;
  }
  tau[N.s + j] = 0.5 * (lo + hi);
  org[N.s + j] = o;
}

// Loewner z-hat (Gu-Eisenstat): zh_i^2 = (lambda_i - d_i) prod_{j != i} (lambda_j - d_i)/(d_j - d_i) / rho.
__global__ void k_dc_zhat(const DcNode* nodes, const int* nk, const double* rho_, const double* kd, const double* kz,
                          const double* tau, const int* org, double* zh) {
  const DcNode N = nodes[blockIdx.y];
  const int K = nk[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K) return;
  const double* d = kd + N.s;
  const double di = d[i];
  double p = ((d[org[N.s + i]] - di) + tau[N.s + i]) / rho_[blockIdx.y];
  for (int j = 0; j < K; j++) {
    if (j == i) continue;
    double lj = (d[org[N.s + j]] - di) + tau[N.s + j];
    p *= lj / (d[j] - di);
  }
  double z = sqrt(fmax(p, 0.0));
  zh[N.s + i] = kz[N.s + i] >= 0.0 ? z : -z;
}

// Eigenvectors of the rank-one problem: U[i][j] = zh_i / (d_i - lambda_j), columns normalised.
// U of node at uoff[node], K x K column-major. One workgroup per column j.
__global__ void k_dc_u(const DcNode* nodes, const int* nk, const long long* uoff, const double* kd, const double* zh,
                       const double* tau, const int* org, double* U) {
  __shared__ double red[16];
  const DcNode N = nodes[blockIdx.y];
  const int K = nk[blockIdx.y];
  const int j = blockIdx.x;
  if (j >= K) return;
  const double* d = kd + N.s;
  const double dj = d[org[N.s + j]], tj = tau[N.s + j];
  double* u = U + uoff[blockIdx.y] + (long long)j * K;
  double s = 0.0;
  for (int i = threadIdx.x; i < K; i += blockDim.x) {
    double x = zh[N.s + i] / ((d[i] - dj) - tj);
    u[i] = x;
    s += x * x;
  }
  s = block_sum(s, red);
  const double inv = 1.0 / sqrt(s);
  for (int i = threadIdx.x; i < K; i += blockDim.x) u[i] *= inv;
}

// Gather the kept columns of Q (node rows) into G (m x K at goff[node]).
__global__ void k_dc_gather(const DcNode* nodes, const int* nk, const long long* goff, const int* kcol, const double* Q,
                            int n, double* G) {
  const DcNode N = nodes[blockIdx.y];
  const int K = nk[blockIdx.y];
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)N.m * K) return;
  const int r = (int)(idx % N.m), k = (int)(idx / N.m);
  G[goff[blockIdx.y] + idx] = Q[(long long)kcol[N.s + k] * n + N.s + r];
}

// New eigenvalues of the node: kept roots lambda_j first, then the deflated poles, with the source
// of each vector (>= 0: column of the GEMM result; < 0: -1 - column of Q). One thread per node.
__global__ void k_dc_vals(const DcNode* nodes, int nnodes, const int* nk, const int* kind, const double* ds,
                          const int* perm, const double* kd, const double* tau, const int* org, double* val, int* src) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nnodes) return;
  const DcNode N = nodes[node];
  const int K = nk[node], m = N.m, s = N.s;
  int k = 0, q = K;
  for (int j = 0; j < m; j++) {
    if (kind[s + j]) {
      val[s + k] = kd[s + org[s + k]] + tau[s + k];
      src[s + k] = k;
      k++;
    } else {
      val[s + q] = ds[s + j];
      src[s + q] = -1 - perm[s + j];
      q++;
    }
  }
}

// Sorted placement: one workgroup per candidate a of a node computes its rank among the node's
// values and copies its vector (GEMM result or deflated Q column) into Q2, the value into D2.
__global__ void k_dc_place(const DcNode* nodes, const long long* goff, const double* val, const int* src,
                           const double* QG, const double* Q, int n, double* Q2, double* D2) {
  __shared__ int rk[4];
  const DcNode N = nodes[blockIdx.y];
  const int a = blockIdx.x, m = N.m, s = N.s;
  if (a >= m) return;
  const double v = val[s + a];
  int c = 0;
  for (int b = threadIdx.x; b < m; b += blockDim.x) {
    double w = val[s + b];
    if (w < v || (w == v && b < a)) c++;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) rk[threadIdx.x >> 6] = c;
  __syncthreads();
  int rank = 0;
  for (int wv = 0; wv < (int)(blockDim.x >> 6); wv++) rank += rk[wv];
  if (threadIdx.x == 0) D2[s + rank] = v;
  const int sc = src[s + a];
  double* dst = Q2 + (long long)(s + rank) * n + s;
  const double* from = sc >= 0 ? QG + goff[blockIdx.y] + (long long)sc * m : Q + (long long)(-1 - sc) * n + s;
  for (int r = threadIdx.x; r < m; r += blockDim.x) dst[r] = from[r];
}

// Copy the merged nodes' blocks (columns and rows [s, s+m)) and eigenvalues back from Q2, D2 into
// Q, D; nodes of other levels keep their blocks. One workgroup per column of a node.
__global__ void k_dc_copyback(const DcNode* nodes, const double* Q2, const double* D2, int n, double* Q, double* D) {
  const DcNode N = nodes[blockIdx.y];
  const int c = blockIdx.x;
  if (c >= N.m) return;
  const long long off = (long long)(N.s + c) * n + N.s;
  for (int r = threadIdx.x; r < N.m; r += blockDim.x) Q[off + r] = Q2[off + r];
  if (threadIdx.x == 0) D[N.s + c] = D2[N.s + c];
}

// ------------------------------------------------------------------ back-transformation
// Explicit panel of reflectors i0..i0+nb-1 (rows i0+1..n-1 of the result): V[r][j] with unit
// diagonal, zero above (reflector j acts on rows i0+j+1..).
__global__ void k_bt_panel(const double* A, int n, int i0, int nb, double* V) {
  const int rows = n - i0 - 1;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)rows * nb) return;
  const int r = (int)(idx % rows), j = (int)(idx / rows);
  const int col = i0 + j;
  // global row of r: i0 + 1 + r; the reflector of column col covers rows col+1..n-1
  const int g = i0 + 1 + r;
  double v = 0.0;
  if (g >= col + 1) v = A[(long long)col * n + g];
  V[idx] = v;
}

// T factor (upper triangular, nb x nb) of the panel from VtV = V^T V and tau (LAPACK dlarft,
// forward columnwise): T[j][j] = tau_j, T[0:j, j] = -tau_j T[0:j, 0:j] VtV[0:j, j]. One workgroup.
__global__ void k_bt_tfactor(const double* VtV, const double* tau, int i0, int nb, double* T) {
  __shared__ double tmp[128];
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) T[e] = 0.0;
  __syncthreads();
  for (int j = 0; j < nb; j++) {
    const double tj = tau[i0 + j];
    if (threadIdx.x < j) {
      int r = threadIdx.x;
      double s = 0.0;
      for (int k = r; k < j; k++) s += T[r + k * nb] * VtV[k + j * nb];
      tmp[r] = -tj * s;
    }
    __syncthreads();
    if (threadIdx.x < j) T[threadIdx.x + j * nb] = tmp[threadIdx.x];
    if (threadIdx.x == 0) T[j + j * nb] = tj;
    __syncthreads();
  }
}

}  // namespace axtb
