// Device code for one shell pair of contracted Gaussians (s, p and d shells): overlap, dipole and
// traceless quadrupole blocks by Obara-Saika 1D moment tables, and the contraction of their
// ket-centre derivatives with the density, energy-weighted density and multipole potentials.
//
// Conventions follow axtb's integrals (crates/axtb/src/integrals.rs:1-15): bra shell cj on R_j,
// ket shell ci on R_i, vec = R_i - R_j, spherical p order (y, z, x), quadrupoles packed
// (xx, xy, yy, xz, yz, zz) in the traceless form 3/2 r_a r_b - 1/2 r^2 delta_ab, moments around
// the ket centre, derivatives with respect to R_i.
#pragma once
#include "common.hpp"

namespace axtb {

// 1D Obara-Saika table t[a][b] = <(x-A)^a (x-B)^b> / (Gaussian prefactor) for a <= LA, b <= LB
// (integrals.rs:103-127).
template <int LA, int LB>
__device__ __forceinline__ void os_table(double rpa, double rpb, double h, double (&t)[LA + 1][LB + 1]) {
  t[0][0] = 1.0;
#pragma unroll
  for (int i = 0; i < LA; i++) t[i + 1][0] = rpa * t[i][0] + (i > 0 ? i * h * t[i - 1][0] : 0.0);
#pragma unroll
  for (int j = 0; j < LB; j++) {
#pragma unroll
    for (int i = 0; i <= LA; i++) {
      double v = rpb * t[i][j];
      if (i > 0) v += i * h * t[i - 1][j];
      if (j > 0) v += j * h * t[i][j - 1];
      t[i][j + 1] = v;
    }
  }
}

__device__ __forceinline__ void traceless6(double* q) {
  double tr = 0.5 * (q[0] + q[2] + q[5]);
  q[0] = 1.5 * q[0] - tr;
  q[1] *= 1.5;
  q[2] = 1.5 * q[2] - tr;
  q[3] *= 1.5;
  q[4] *= 1.5;
  q[5] = 1.5 * q[5] - tr;
}

// Moments around R_j -> around R_i given vec = R_i - R_j ... and the reverse shift used for the
// transposed element: moments around the ket centre -> around the bra centre
// (integrals.rs:991-1006): D_j = D_i + vec S, Q_j = traceless(shift) + Q_i.
__device__ __forceinline__ void shift_to_bra(const double* vec, double s, const double* di, const double* qi,
                                             double* dj, double* qj) {
  dj[0] = di[0] + vec[0] * s;
  dj[1] = di[1] + vec[1] * s;
  dj[2] = di[2] + vec[2] * s;
  double t[6] = {2.0 * vec[0] * di[0] + vec[0] * vec[0] * s, vec[0] * di[1] + vec[1] * di[0] + vec[0] * vec[1] * s,
                 2.0 * vec[1] * di[1] + vec[1] * vec[1] * s, vec[0] * di[2] + vec[2] * di[0] + vec[0] * vec[2] * s,
                 vec[1] * di[2] + vec[2] * di[1] + vec[1] * vec[2] * s, 2.0 * vec[2] * di[2] + vec[2] * vec[2] * s};
  traceless6(t);
#pragma unroll
  for (int k = 0; k < 6; k++) qj[k] = t[k] + qi[k];
}

// Cartesian components of a shell in CCA order (integrals.rs:29-45): s; p x, y, z; d xx, xy, xz,
// yy, yz, zz. cart_exp(l, c, k) is the exponent of axis k in component c.
__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int cart_exp(int l, int c, int k) {
  return l == 0   ? 0
         : l == 1 ? (c == k ? 1 : 0)
                  : (k == 0   ? (c == 0 ? 2 : (c == 1 || c == 2) ? 1 : 0)
                     : k == 1 ? (c == 3 ? 2 : (c == 1 || c == 4) ? 1 : 0)
                              : (c == 5 ? 2 : (c == 2 || c == 4) ? 1 : 0));
}

// Weight of Cartesian component c in real spherical function m of shell l (integrals.rs:49-103):
// p (y, z, x); d (sqrt3 xy, sqrt3 yz, zz - (xx + yy) / 2, sqrt3 xz, sqrt3 / 2 (xx - yy)), the
// Cartesian primitives normalised with tblite's 1 / sqrt((2l - 1)!!).
__host__ __device__ constexpr double sph_coef(int l, int m, int c) {
  constexpr double S3 = 1.7320508075688772;  // sqrt(3)
  return l == 0   ? 1.0
         : l == 1 ? (((m == 0 && c == 1) || (m == 1 && c == 2) || (m == 2 && c == 0)) ? 1.0 : 0.0)
         : m == 0 ? (c == 1 ? S3 : 0.0)
         : m == 1 ? (c == 4 ? S3 : 0.0)
         : m == 2 ? ((c == 0 || c == 3) ? -0.5 : (c == 5 ? 1.0 : 0.0))
         : m == 3 ? (c == 2 ? S3 : 0.0)
                  : (c == 0 ? 0.5 * S3 : (c == 3 ? -0.5 * S3 : 0.0));
}

// S, D (3), Q (6) of the ket rows [IM0, IM0 + NR) of a shell pair with the angular momenta fixed at
// compile time (blocks in registers): out[(im - IM0) * NSJ + jm][10] = {s, dx, dy, dz, qxx, qxy,
// qyy, qxz, qyz, qzz}. Per primitive pair the Cartesian moments of every component pair are
// formed and added into the spherical blocks that use them (weights known at compile time; the
// weight 1 of s and p adds the moment itself).
template <int LI, int LJ, int IM0, int NR>
__device__ __forceinline__ void shellpair_multipole_t(int npi, const double* ai_, const double* ci_, int npj,
                                                      const double* aj_, const double* cj_, const double* vec,
                                                      double intcut, double (&out)[NR * (2 * LJ + 1)][10]) {
  constexpr int NSJ = 2 * LJ + 1, NCI = ncart(LI), NCJ = ncart(LJ);
  const double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
#pragma unroll
  for (int b = 0; b < NR * NSJ; b++)
#pragma unroll
    for (int x = 0; x < 10; x++) out[b][x] = 0.0;
  const double sqrtpi3 = 5.568327996831708;
  for (int ip = 0; ip < npi; ip++) {
    const double ai = ai_[ip];
    for (int jp = 0; jp < npj; jp++) {
      const double aj = aj_[jp];
      const double eab = ai + aj, oab = 1.0 / eab, est = ai * aj * r2 * oab;
      if (est > intcut) continue;
      const double so = sqrt(oab);
      const double cc = ci_[ip] * cj_[jp] * (exp(-est) * sqrtpi3 * (so * so * so));
      double t[3][LI + 3][LJ + 1];
#pragma unroll
      for (int k = 0; k < 3; k++) os_table<LI + 2, LJ>(-vec[k] * aj * oab, vec[k] * ai * oab, 0.5 * oab, t[k]);
#pragma unroll
      for (int ic = 0; ic < NCI; ic++) {
        bool used_i = false;
#pragma unroll
        for (int r = 0; r < NR; r++) used_i |= sph_coef(LI, IM0 + r, ic) != 0.0;
        if (!used_i) continue;
#pragma unroll
        for (int jc = 0; jc < NCJ; jc++) {
          double v[3][3];
#pragma unroll
          for (int k = 0; k < 3; k++) {
            const int a = cart_exp(LI, ic, k), b = cart_exp(LJ, jc, k);
#pragma unroll
            for (int m = 0; m < 3; m++) v[k][m] = t[k][a + m][b];
          }
          double val[10];
          val[0] = cc * v[0][0] * v[1][0] * v[2][0];
          val[1] = cc * v[0][1] * v[1][0] * v[2][0];
          val[2] = cc * v[0][0] * v[1][1] * v[2][0];
          val[3] = cc * v[0][0] * v[1][0] * v[2][1];
          val[4] = cc * v[0][2] * v[1][0] * v[2][0];
          val[5] = cc * v[0][1] * v[1][1] * v[2][0];
          val[6] = cc * v[0][0] * v[1][2] * v[2][0];
          val[7] = cc * v[0][1] * v[1][0] * v[2][1];
          val[8] = cc * v[0][0] * v[1][1] * v[2][1];
          val[9] = cc * v[0][0] * v[1][0] * v[2][2];
#pragma unroll
          for (int r = 0; r < NR; r++) {
            const double wi = sph_coef(LI, IM0 + r, ic);
            if (wi == 0.0) continue;
#pragma unroll
            for (int jm = 0; jm < NSJ; jm++) {
              const double wj = sph_coef(LJ, jm, jc);
              if (wj == 0.0) continue;
              const double w = wi * wj;
              double* o = out[r * NSJ + jm];
              if (w == 1.0) {
#pragma unroll
                for (int x = 0; x < 10; x++) o[x] += val[x];
              } else {
#pragma unroll
                for (int x = 0; x < 10; x++) o[x] += w * val[x];
              }
            }
          }
        }
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NR * NSJ; b++) traceless6(&out[b][4]);
}

// Contraction of the ket-centre derivatives of S, D and Q of a shell pair with the block
// coefficients and the multipole potentials of the two atoms (hamiltonian.rs:169-211):
//   dE/dR_i += sum_blk [ sval ds + hp2 s dsv - p (ddi.vdp_i + ddj.vdp_j + dqi.vqp_i + dqj.vqp_j) ];
// returns sum_blk p s (for dE/dCN). The derivatives are those of integrals.rs:400-521, contracted
// per primitive pair (all terms are linear in the primitive contributions), for the ket rows
// [IM0, IM0 + NR):
// sval, hp2, pp hold the coefficients of those rows' blocks (index (im - IM0) * NSJ + jm). The
// contraction is linear in the block coefficients, so per Cartesian component pair the spherical
// coefficients are folded into Cartesian ones (weights known at compile time; for s and p the
// coefficient itself) and contracted with the Cartesian value and derivative moments.
template <int LI, int LJ, int IM0, int NR>
__device__ __forceinline__ double shellpair_grad_t(int npi, const double* ai_, const double* ci_, int npj,
                                                   const double* aj_, const double* cj_, const double* vec,
                                                   double intcut, const double* sval, const double* hp2,
                                                   const double* pp, const double* dsv, const double* vdi,
                                                   const double* vdj, const double* vqi, const double* vqj,
                                                   double* dg) {
  constexpr int NSJ = 2 * LJ + 1, NCI = ncart(LI), NCJ = ncart(LJ);
  const double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
  const double sqrtpi3 = 5.568327996831708;
  double ps = 0.0;
  for (int ip = 0; ip < npi; ip++) {
    const double ai = ai_[ip];
    const double ai2 = 2.0 * ai;
    for (int jp = 0; jp < npj; jp++) {
      const double aj = aj_[jp];
      const double eab = ai + aj, oab = 1.0 / eab, est = ai * aj * r2 * oab;
      if (est > intcut) continue;
      const double so = sqrt(oab);
      const double cc = ci_[ip] * cj_[jp] * (exp(-est) * sqrtpi3 * (so * so * so));
      double t[3][LI + 2][LJ + 3];
#pragma unroll
      for (int k = 0; k < 3; k++) os_table<LI + 1, LJ + 2>(-vec[k] * aj * oab, vec[k] * ai * oab, 0.5 * oab, t[k]);
#pragma unroll
      for (int ic = 0; ic < NCI; ic++) {
        bool used_i = false;
#pragma unroll
        for (int r = 0; r < NR; r++) used_i |= sph_coef(LI, IM0 + r, ic) != 0.0;
        if (!used_i) continue;
#pragma unroll
        for (int jc = 0; jc < NCJ; jc++) {
          // the Cartesian coefficients of this component pair
          double csv = 0.0, chp = 0.0, cp = 0.0;
          bool used = false;
#pragma unroll
          for (int r = 0; r < NR; r++) {
            const double wi = sph_coef(LI, IM0 + r, ic);
            if (wi == 0.0) continue;
#pragma unroll
            for (int jm = 0; jm < NSJ; jm++) {
              const double wj = sph_coef(LJ, jm, jc);
              if (wj == 0.0) continue;
              const double w = wi * wj;
              const int blk = r * NSJ + jm;
              if (!used && w == 1.0) {
                csv = sval[blk];
                chp = hp2[blk];
                cp = pp[blk];
              } else {
                csv += w * sval[blk];
                chp += w * hp2[blk];
                cp += w * pp[blk];
              }
              used = true;
            }
          }
          if (!used) continue;
          // moments around R_j and their R_i derivatives, per direction
          double v[3][3], g[3][3];
#pragma unroll
          for (int k = 0; k < 3; k++) {
            const int a = cart_exp(LI, ic, k), b = cart_exp(LJ, jc, k);
#pragma unroll
            for (int m = 0; m < 3; m++) {
              v[k][m] = t[k][a][b + m];
              g[k][m] = ai2 * t[k][a + 1][b + m] - (a > 1 ? a * t[k][a - 1][b + m] : (a > 0 ? t[k][0][b + m] : 0.0));
            }
          }
          constexpr int MOM[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0},
                                      {1, 1, 0}, {0, 2, 0}, {1, 0, 1}, {0, 1, 1}, {0, 0, 2}};
          double val[10], der[10][3];
#pragma unroll
          for (int n = 0; n < 10; n++) {
            const double x = v[0][MOM[n][0]], y = v[1][MOM[n][1]], z = v[2][MOM[n][2]];
            val[n] = cc * x * y * z;
            der[n][0] = cc * g[0][MOM[n][0]] * y * z;
            der[n][1] = cc * x * g[1][MOM[n][1]] * z;
            der[n][2] = cc * x * y * g[2][MOM[n][2]];
          }
          const double s = val[0];
          const double pij = cp;
          ps += pij * s;
#pragma unroll
          for (int k = 0; k < 3; k++) {
            const double ds = der[0][k];
            double ddj[3], ddi[3], dmj[6], dmi[6];
#pragma unroll
            for (int c = 0; c < 3; c++) {
              ddj[c] = der[1 + c][k];
              ddi[c] = ddj[c] - vec[c] * ds;
            }
            ddi[k] -= s;
            constexpr int AB[6][2] = {{0, 0}, {0, 1}, {1, 1}, {0, 2}, {1, 2}, {2, 2}};
#pragma unroll
            for (int n = 0; n < 6; n++) {
              const int a = AB[n][0], b = AB[n][1];
              dmj[n] = der[4 + n][k];
              double x = dmj[n] - vec[a] * ddj[b] - vec[b] * ddj[a] + vec[a] * vec[b] * ds;
              if (a == k) x += -val[1 + b] + vec[b] * s;
              if (b == k) x += -val[1 + a] + vec[a] * s;
              dmi[n] = x;
            }
            traceless6(dmi);
            traceless6(dmj);
            double mp = 0.0;
#pragma unroll
            for (int c = 0; c < 3; c++) mp += ddi[c] * vdi[c] + ddj[c] * vdj[c];
#pragma unroll
            for (int c = 0; c < 6; c++) mp += dmi[c] * vqi[c] + dmj[c] * vqj[c];
            dg[k] += csv * ds + chp * s * dsv[k] - pij * mp;
          }
        }
      }
    }
  }
  return ps;
}

}  // namespace axtb
