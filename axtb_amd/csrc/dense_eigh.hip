// Dense symmetric eigensolver for large matrices (FP64): Householder reduction to tridiagonal
// form, divide and conquer on the tridiagonal matrix (Cuppen's tearing, Gu-Eisenstat eigenvectors
// with deflation), and the back-transformation of the eigenvectors with the reflectors in blocked
// (WY) form on the matrix cores. All matrices column-major.
#include "common.hpp"

namespace axtb {

// ------------------------------------------------------------------ Householder tridiagonalisation
// Blocked reduction to tridiagonal form (LAPACK dsytrd/dlatrd, lower): panels of TRD_NB columns;
// within a panel the trailing matrix is left as it was and corrected through the panel's
// reflectors V (kept in A below the diagonal, unit first element stored) and W; at the end of a
// panel A22 -= V W^T + W V^T runs as one two-term GEMM on the matrix cores. Two launches per column:
// k_trd_col (W column of the previous reflector, update of the column, new reflector) and
// k_trd_symv (A22 v in row-block x column-block partials from the lower triangle, W^T v, V^T v and
// v^T A22 v partials).
// With y1 = W^T v, y2 = V^T v and p = A22 v of reflector v (scalar tau):
//   w = tau (p - V y1 - W y2) + alpha v,  alpha = -tau^2 (v^T p - 2 y1.y2) / 2.
// Only the lower triangle of A is read and kept up to date (the trailing update skips the tiles
// above the diagonal).
constexpr int TRD_NB = 64;   // panel width
constexpr int TRD_CB = 64;   // rows per workgroup of the column kernel (4 wavefronts share the terms)
constexpr int TRD_CH = 64;   // symv tile (TRD_CH x TRD_CH of the lower triangle)
constexpr int TRD_DR = 1024; // rows per dot workgroup (one column per wavefront)

// Reflector of x = (ax, x1...) with sum(x1^2) = tot: H x = (beta, 0, ...), v = (1, x1 * scal).
__device__ __forceinline__ void trd_reflector(double ax, double tot, double& tau, double& beta, double& scal) {
  if (tot == 0.0) {
    tau = 0.0;
    beta = ax;
    scal = 1.0;
    return;
  }
  const double nrm = sqrt(ax * ax + tot);
  beta = ax >= 0.0 ? -nrm : nrm;
  tau = (beta - ax) / beta;
  scal = 1.0 / (ax - beta);
}

// s = sum of p[k * stride] over k = k0, k0 + step, ... < n, in that order, with the loads issued 8 at
// a time (a plain loop waits for each load before it issues the next).
__device__ __forceinline__ double strided_sum(const double* p, int k0, int n, int step, long long stride) {
  double s = 0.0;
  for (int k = k0; k < n; k += 8 * step) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int kk = k + u * step;
      v[u] = kk < n ? p[(long long)kk * stride] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (k + u * step < n) s += v[u];
  }
  return s;
}

// Sum of the column kernel's squares partials of column i (fixed order; the same in every lane
// and every wavefront).
__device__ __forceinline__ double trd_sumsq(const double* sq, int n, int i, int lane) {
  const int nq = (n - i + TRD_CB - 1) / TRD_CB;
  double t = 0.0;
  for (int q = lane; q < nq; q += 64) t += sq[q];
  return wave_sum(t);
}

// Column kernel for column i = i0 + j, rows [i, n), 64 rows per workgroup of 4 wavefronts (the
// wavefronts share the column terms, c = w mod 4). For the previous column ip = i - 1 (j > 0):
// its reflector (from the squares partials and the stored first element ax[ip]) is scaled in
// place, tau/e/d written, and W(:, j-1) formed from the symv partials. Then (unless
// finalize_only) column i is updated with the panel's reflectors so far, its first
// sub-diagonal element saved in ax[i] and its squares partials written to sq_cur. Nothing here
// reads what another workgroup of the same launch writes.
__global__ __launch_bounds__(256) void k_trd_col(double* A, int n, int i0, int j, double* W, const double* part,
                                                 const double* dpart, const double* vpart, const double* sq_prev,
                                                 double* sq_cur, double* axv, double* tau, double* d, double* e,
                                                 int finalize_only) {
  __shared__ double y1[TRD_NB], y2[TRD_NB], wi[TRD_NB], vi[TRD_NB];
  __shared__ double red[3][4][64];
  __shared__ double rs[16];
  const int i = i0 + j, jj = j - 1, ip = i - 1;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = i + blockIdx.x * TRD_CB + lane;  // global row
  const long long ln = n;
  const int mp = n - i;  // rows of the previous reflector and of its symv
  const int nch = j > 0 ? (mp + TRD_CH - 1) / TRD_CH : 0;
  // this row's loads first (none depends on the reductions below, so they overlap them): its
  // share of the symv partials, V(r, c) and W(r, c) for c = wv + 4k < jj, and A(r, ip), A(r, i)
  constexpr int KR = (TRD_NB + 3) / 4;
  double ps = 0.0, vr[KR], wr_[KR], a_ip = 0.0, a_i = 0.0;
  if (r < n) {
    ps = strided_sum(part + r, wv, nch, 4, ln);
#pragma unroll
    for (int k = 0; k < KR; k++) {
      const int c = wv + 4 * k;
      vr[k] = c < jj ? A[(i0 + c) * ln + r] : 0.0;
      wr_[k] = c < jj ? W[c * ln + r] : 0.0;
    }
    if (wv == 0) {
      if (j > 0) a_ip = A[ip * ln + r];
      a_i = A[i * ln + r];
    }
  }
  // row i's coefficients V(i, c), W(i, c) (thread c) and wave 0's share of row i's symv partials,
  // issued with the loads above
  double vic = 0.0, wic = 0.0, pin = 0.0;
  if (!finalize_only) {
    if ((int)threadIdx.x < jj) {
      vic = A[(i0 + threadIdx.x) * ln + i];
      wic = W[threadIdx.x * ln + i];
    }
    if (wv == 0) pin = strided_sum(part + i, lane, nch, 64, ln);
  }
  double t = 0.0, alpha = 0.0, scal = 1.0;
  if (j > 0) {
    const int ndb = (mp + TRD_DR - 1) / TRD_DR;
    const int ntile = nch * (nch + 1) / 2;
    double beta;
    trd_reflector(axv[ip], trd_sumsq(sq_prev, n, ip, lane), t, beta, scal);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      tau[ip] = t;
      e[ip] = beta;
      d[ip] = A[ip * ln + ip];
    }
    for (int c = threadIdx.x; c < 2 * jj; c += 256) {
      const double s = strided_sum(dpart + c, 0, ndb, 1, 2 * TRD_NB);
      if (c < jj)
        y1[c] = s;
      else
        y2[c - jj] = s;
    }
    double va = strided_sum(vpart, threadIdx.x, ntile, 256, 1);
    va = block_sum(va, rs);  // its barriers also publish y1, y2
    const double yy = wave_sum(lane < jj ? y1[lane] * y2[lane] : 0.0);
    alpha = -0.5 * t * t * (va - 2.0 * yy);
  }
  if (!finalize_only) {
    // coefficients of row i: V(i, c), W(i, c); V(i, jj) is the unit first element
    const int c = threadIdx.x;
    if (c < jj) {
      vi[c] = vic;
      wi[c] = wic;
    }
    if (j > 0 && wv == 0) {
      const double p = wave_sum(pin);
      const double sw = wave_sum(lane < jj ? vic * y1[lane] + wic * y2[lane] : 0.0);
      if (lane == 0) {
        vi[jj] = 1.0;
        wi[jj] = t * (p - sw) + alpha;
      }
    }
    __syncthreads();
  }
  // per wavefront: its share of sw_r = V y1 + W y2 and of u_r = V wi + W vi
  double sw = 0.0, u = 0.0;
  if (r < n) {
#pragma unroll
    for (int k = 0; k < KR; k++) {
      const int c = wv + 4 * k;
      if (c < jj) {
        sw += vr[k] * y1[c] + wr_[k] * y2[c];
        u += vr[k] * wi[c] + wr_[k] * vi[c];
      }
    }
  }
  red[0][wv][lane] = ps;
  red[1][wv][lane] = sw;
  red[2][wv][lane] = u;
  __syncthreads();
  if (wv != 0) return;
  double a = 0.0;
  if (r < n) {
    const double p = red[0][0][lane] + red[0][1][lane] + red[0][2][lane] + red[0][3][lane];
    const double swt = red[1][0][lane] + red[1][1][lane] + red[1][2][lane] + red[1][3][lane];
    double ut = red[2][0][lane] + red[2][1][lane] + red[2][2][lane] + red[2][3][lane];
    if (j > 0) {
      const double vp = r == i ? 1.0 : a_ip * scal;
      A[ip * ln + r] = vp;
      const double wr = t * (p - swt) + alpha * vp;
      W[jj * ln + r] = wr;
      if (!finalize_only) ut += vp * wi[jj] + wr * vi[jj];
    }
    if (!finalize_only) {
      a = a_i - ut;
      A[i * ln + r] = a;
      if (r == i + 1) axv[i] = a;
    }
  }
  if (finalize_only) return;
  const double s = wave_sum((r < n && r >= i + 2) ? a * a : 0.0);
  if (lane == 0) sq_cur[blockIdx.x] = s;
}

// Partials of p = A22 v (A22 = A[i+1:, i+1:], v the reflector of column i, scaled on the fly)
// from the lower triangle only: the first ntw workgroups take the nt = nb (nb + 1) / 2 64 x 64 tiles
// (rb, cb), rb >= cb, of A22 (nb = ceil(m / 64)), tpw consecutive tiles each, the next tile's loads
// issued before the current one is summed; tile (rb, cb) writes the row sums of its rows as
// partial cb (part[cb][rows of rb]) and, below the diagonal, its column sums A^T v as partial rb of
// the rows of block cb (part[rb][rows of cb]), so that every row r gets exactly one partial q for
// each q < nb; vpart[t] = v^T (both sums) of tile t. Diagonal tiles read the upper half from its
// mirror in the lower half. The workgroups after ntw form W^T v and V^T v partials, one column per
// wavefront over 1024 rows.
__global__ __launch_bounds__(256) void k_trd_symv(const double* A, int n, int i0, int j, const double* W,
                                                  const double* sq, const double* axv, double* part, double* dpart,
                                                  double* vpart, int tpw) {
  __shared__ double tile[TRD_CH][TRD_CH + 1];  // tile[c][r] = A22(R0 + r, C0 + c)
  __shared__ double vr[TRD_CH], vc[TRD_CH];
  __shared__ double red[2][4][TRD_CH];
  const int i = i0 + j, m = n - i - 1;
  const long long ln = n;
  const double* x = A + i * ln + i + 1;
  const int nb = (m + TRD_CH - 1) / TRD_CH, nt = nb * (nb + 1) / 2, ndb = (m + TRD_DR - 1) / TRD_DR;
  const int ntw = (nt + tpw - 1) / tpw;  // tile workgroups, tpw consecutive tiles each
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double t, beta, scal;
  trd_reflector(axv[i], trd_sumsq(sq, n, i, lane), t, beta, scal);
  auto vec = [&](int k) { return k == 0 ? 1.0 : x[k] * scal; };
  if ((int)blockIdx.x < ntw) {
    constexpr int KT = TRD_CH * TRD_CH / 256;
    auto tile_of = [](int q, int& rb, int& cb) {
      rb = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
      while ((rb + 1) * (rb + 2) / 2 <= q) rb++;
      while (rb * (rb + 1) / 2 > q) rb--;
      cb = q - rb * (rb + 1) / 2;
    };
    // all 16 loads of this thread for tile q in flight at once (lane = row, columns wv + 4 k)
    auto load = [&](int q, double (&ld)[KT]) {
      int rb, cb;
      tile_of(q, rb, cb);
      const int R0 = rb * TRD_CH, C0 = cb * TRD_CH;
      const double* a = A + (long long)(i + 1 + C0) * ln + (i + 1 + R0);
      const bool rin = R0 + lane < m;
#pragma unroll
      for (int k = 0; k < KT; k++) {
        const int c = wv + 4 * k;
        ld[k] = (rin && C0 + c < m) ? a[c * ln + lane] : 0.0;
      }
    };
    const int qa = blockIdx.x * tpw, qe = min(qa + tpw, nt);
    double ld[KT], nx[KT];
    load(qa, ld);
    for (int q = qa; q < qe; q++) {
      // the next tile's loads are in flight while this one is summed
      if (q + 1 < qe) load(q + 1, nx);
      if (q > qa) __syncthreads();  // the previous tile's reads of tile, vr, vc are done
      int rb, cb;
      tile_of(q, rb, cb);
      const int R0 = rb * TRD_CH, C0 = cb * TRD_CH;
#pragma unroll
      for (int k = 0; k < KT; k++) tile[wv + 4 * k][lane] = ld[k];
      if (tid < TRD_CH) {
        vr[tid] = R0 + tid < m ? vec(R0 + tid) : 0.0;
      } else if (tid < 2 * TRD_CH) {
        const int c = tid - TRD_CH;
        vc[c] = C0 + c < m ? vec(C0 + c) : 0.0;
      }
      __syncthreads();
      const bool diag = rb == cb;
      // row sums: lane = row, wavefront = quarter of the columns
      double s = 0.0, u = 0.0;
      const int k0 = wv * (TRD_CH / 4);
#pragma unroll 8
      for (int c = k0; c < k0 + TRD_CH / 4; c++) s += (diag && c > lane ? tile[lane][c] : tile[c][lane]) * vc[c];
      // column sums (below the diagonal): lane = column, wavefront = quarter of the rows
      if (!diag) {
#pragma unroll 8
        for (int r = k0; r < k0 + TRD_CH / 4; r++) u += tile[lane][r] * vr[r];
      }
      red[0][wv][lane] = s;
      red[1][wv][lane] = u;
      __syncthreads();
      if (wv == 0) {
        const double rs = red[0][0][lane] + red[0][1][lane] + red[0][2][lane] + red[0][3][lane];
        if (R0 + lane < m) part[cb * ln + (i + 1 + R0 + lane)] = rs;
        double pv = rs * vr[lane];
        if (!diag) {
          const double cs = red[1][0][lane] + red[1][1][lane] + red[1][2][lane] + red[1][3][lane];
          if (C0 + lane < m) part[rb * ln + (i + 1 + C0 + lane)] = cs;
          pv += cs * vc[lane];
        }
        pv = wave_sum(pv);
        if (lane == 0) vpart[q] = pv;
      }
      if (q + 1 < qe) {
#pragma unroll
        for (int k = 0; k < KT; k++) ld[k] = nx[k];
      }
    }
  } else {
    const int d = blockIdx.x - ntw;
    const int rb = d % ndb;
    const int c = 4 * (d / ndb) + wv;
    if (c < 2 * j) {
      const double* y = c < j ? W + c * ln : A + (i0 + c - j) * ln;
      // all 2 x 16 loads of the lane first, then the products in row order
      constexpr int NQ = TRD_DR / 64;
      double yv[NQ], xv[NQ];
#pragma unroll
      for (int u = 0; u < NQ; u++) {
        const int rl = rb * TRD_DR + lane + 64 * u;
        yv[u] = rl < m ? y[i + 1 + rl] : 0.0;
        xv[u] = rl < m ? x[rl] : 0.0;
      }
      double s = 0.0;
#pragma unroll
      for (int u = 0; u < NQ; u++) {
        const int rl = rb * TRD_DR + lane + 64 * u;
        if (rl < m) s += yv[u] * (rl == 0 ? 1.0 : xv[u] * scal);
      }
      s = wave_sum(s);
      if (lane == 0) dpart[rb * 2 * TRD_NB + c] = s;
    }
  }
}

// The last diagonal element (no reflector for the last column) and zero padding of e, tau.
__global__ void k_hh_last(const double* A, int n, double* d, double* e, double* tau) {
  d[n - 1] = A[(long long)(n - 1) * n + (n - 1)];
  tau[n - 1] = 0.0;
  e[n - 1] = 0.0;
}

__global__ void k_zero_u32(unsigned* p, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0u;
}

// ------------------------------------------------------------------ blocked Cholesky start basis
// Diagonal block of the right-looking blocked Cholesky factorisation S = L L^T (the reduction of
// GenEigSolver, base/src/linalg/eigen.rs:72-113): the b x b block at (k0, k0) of W (lower
// triangle, already updated by the previous panels) is factored in LDS and written back, and its
// inverse L11^-1 (lower, zero above) is written to M at (k0, k0). fail = 1 on a pivot <= 0.
__global__ __launch_bounds__(256) void k_chol_diag(double* W, int n, int k0, int b, double* M, int* fail) {
  __shared__ double a[64][65];  // a[c][r]
  __shared__ double mi[64][65];
  __shared__ int bad;
  const int tid = threadIdx.x;
  const long long ln = n;
  if (tid == 0) bad = 0;
  for (int e = tid; e < 64 * 64; e += 256) {
    const int r = e & 63, c = e >> 6;
    a[c][r] = (r < b && c < b && r >= c) ? W[(k0 + r) + (k0 + c) * ln] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < b; j++) {
    if (tid == 0) {
      const double d = a[j][j];
      if (!(d > 0.0)) bad = 1;
      a[j][j] = d > 0.0 ? sqrt(d) : 1.0;
    }
    __syncthreads();
    const double dj = a[j][j];
    for (int r = j + 1 + tid; r < b; r += 256) a[j][r] /= dj;
    __syncthreads();
    // trailing lower block: a(r, c) -= a(r, j) a(c, j), j < c <= r
    const int m = b - j - 1;
    for (int e = tid; e < m * m; e += 256) {
      const int r = j + 1 + e % m, c = j + 1 + e / m;
      if (r >= c) a[c][r] -= a[j][r] * a[j][c];
    }
    __syncthreads();
  }
  // inverse by forward substitution, one column per thread: L x = e_c
  if (tid < b) {
    const int c = tid;
    for (int r = 0; r < b; r++) {
      double v = 0.0;
      if (r == c) v = 1.0;
      if (r >= c) {
        for (int k = c; k < r; k++) v -= a[k][r] * mi[c][k];
        v /= a[r][r];
      }
      mi[c][r] = v;
    }
  }
  __syncthreads();
  for (int e = tid; e < b * b; e += 256) {
    const int r = e % b, c = e / b;
    if (r >= c) W[(k0 + r) + (k0 + c) * ln] = a[c][r];
    M[(k0 + r) + (k0 + c) * ln] = r >= c ? mi[c][r] : 0.0;
  }
  if (tid == 0 && bad) *fail = 1;
}

// X = M^T for lower triangular M (X upper triangular, zero below), by 32 x 32 tiles through LDS.
__global__ __launch_bounds__(256) void k_lower_to_upper_t(const double* M, double* X, int n) {
  __shared__ double tl[32][33];
  const int R0 = blockIdx.y * 32, C0 = blockIdx.x * 32;  // tile of X
  const long long ln = n;
  // X(R0 + i, C0 + j) = M(C0 + j, R0 + i): read M's tile (rows C0.., cols R0..) coalesced
  for (int e = threadIdx.x; e < 32 * 32; e += 256) {
    const int r = e & 31, c = e >> 5;  // M(C0 + r, R0 + c)
    const int gr = C0 + r, gc = R0 + c;
    tl[c][r] = (gr < n && gc < n && gr >= gc) ? M[gr + gc * ln] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 32 * 32; e += 256) {
    const int i = e & 31, j = e >> 5;  // X(R0 + i, C0 + j) = M(C0 + j, R0 + i) = tl[i][j]
    const int gr = R0 + i, gc = C0 + j;
    if (gr < n && gc < n) X[gr + gc * ln] = tl[i][j];
  }
}

// Upper triangle of the n x n matrix A (leading dimension ld) from its lower triangle, by 32 x 32
// tiles through LDS (workgroup t takes lower tile (rb, cb), t = rb (rb + 1) / 2 + cb).
__global__ __launch_bounds__(256) void k_mirror_lower(double* A, int n, long long ld) {
  __shared__ double tl[32][33];
  const int q = blockIdx.x;
  int rb = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
  while ((rb + 1) * (rb + 2) / 2 <= q) rb++;
  while (rb * (rb + 1) / 2 > q) rb--;
  const int cb = q - rb * (rb + 1) / 2;
  const int R0 = rb * 32, C0 = cb * 32;
  for (int e = threadIdx.x; e < 32 * 32; e += 256) {
    const int r = e & 31, c = e >> 5;
    if (R0 + r < n && C0 + c < n) tl[c][r] = A[(R0 + r) + (C0 + c) * ld];
  }
  __syncthreads();
  // element (C0 + c', R0 + r') of the upper triangle = (R0 + r', C0 + c'); consecutive threads
  // write consecutive rows C0 + c'
  for (int e = threadIdx.x; e < 32 * 32; e += 256) {
    const int cp = e & 31, rp = e >> 5;
    const int gr = C0 + cp, gc = R0 + rp;
    if (gr < n && gc < n && gc > gr) A[gr + gc * ld] = tl[cp][rp];
  }
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
// relative to the nearer pole (org), which keeps lambda - d_i accurate. One wavefront per root
// (4 per workgroup), the lanes sharing the sum; the butterfly sum is identical in every lane,
// so the bisection stays uniform across the wavefront.
__global__ __launch_bounds__(256) void k_dc_secular(const DcNode* nodes, const int* nk, const double* rho_,
                                                    const double* kd, const double* kz, double* tau, int* org) {
  const DcNode N = nodes[blockIdx.y];
  const int K = nk[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= K) return;
  const double rho = rho_[blockIdx.y];
  const double* d = kd + N.s;
  const double* z = kz + N.s;
  auto secular = [&](double dorg, double t) {
    double g = 0.0;
    for (int i = lane; i < K; i += 64) g += z[i] * z[i] / ((d[i] - dorg) - t);
    return 1.0 / rho + wave_sum(g);
  };
  const double lo_d = d[j];
  const double hi_d = j + 1 < K ? d[j + 1] : d[j] + rho;
  // the sign at the midpoint decides the origin
  const double mid = 0.5 * (hi_d - lo_d);
  const double f = secular(lo_d, mid);
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
    const double t = 0.5 * (lo + hi);
    if (t == lo || t == hi) break;
    if (secular(dorg, t) > 0.0)
      hi = t;
    else
      lo = t;
  }
  if (lane == 0) {
    tau[N.s + j] = 0.5 * (lo + hi);
    org[N.s + j] = o;
  }
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
// first element, zero above; a reflector with tau = 0 (the identity) gets a zero column.
__global__ void k_bt_panel(const double* A, const double* tau, int n, int i0, int nb, double* V) {
  const int rows = n - i0 - 1;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)rows * nb) return;
  const int r = (int)(idx % rows), j = (int)(idx / rows);
  const int col = i0 + j;
  // global row of r: i0 + 1 + r; the reflector of column col covers rows col+1..n-1
  const int g = i0 + 1 + r;
  double v = 0.0;
  if (g >= col + 1 && tau[col] != 0.0) v = A[(long long)col * n + g];
  V[idx] = v;
}

// The panel's product H_i0 ... H_i0+nb-1 = I - V T V^T has T^-1 = striu(V^T V) + diag(1/tau)
// (the UT transform; an identity reflector has a zero column and the diagonal entry 1). T, upper
// triangular, by inverting T^-1 column by column (one thread per column of T) from LDS; VtV is
// overwritten by T.
__global__ __launch_bounds__(64) void k_bt_tfactor(double* VtV, const double* tau, int i0, int nb) {
  __shared__ double L[64][65];  // T^-1, then T
  __shared__ double X[64][65];
  const int c = threadIdx.x;
  for (int r = 0; r < nb; r++) {
    double v = 0.0;
    if (c < nb) {
      const double t = tau[i0 + r];
      v = c > r ? VtV[r + c * nb] : (c == r ? (t != 0.0 ? 1.0 / t : 1.0) : 0.0);
    }
    L[r][c] = v;
  }
  __syncthreads();
  // column c of T: solve T^-1 x = e_c (back substitution, rows c..0)
  if (c < nb) {
    for (int r = nb - 1; r > c; r--) X[r][c] = 0.0;
    for (int r = c; r >= 0; r--) {
      double acc = r == c ? 1.0 : 0.0;
      for (int q = r + 1; q <= c; q++) acc -= L[r][q] * X[q][c];
      X[r][c] = tau[i0 + r] != 0.0 ? acc * tau[i0 + r] : acc;
    }
  }
  __syncthreads();
  if (c < nb)
    for (int r = 0; r < nb; r++) VtV[r + c * nb] = X[r][c];
}

}  // namespace axtb
