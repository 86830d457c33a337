// Device code for one shell pair of contracted Gaussians (s and p shells): overlap, dipole and
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

// Cartesian component (x=0, y=1, z=2) of spherical p function m (y, z, x); s: -1.
__device__ __forceinline__ int p_cart(int l, int m) { return l == 0 ? -1 : (m == 0 ? 1 : (m == 1 ? 2 : 0)); }

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

// S, D (3), Q (6) of the block [jm + nsj*im] of a shell pair with both l <= 1 (integrals.rs:277-367).
// out[blk][10] = {s, dx, dy, dz, qxx, qxy, qyy, qxz, qyz, qzz}.
__device__ void shellpair_multipole(int li, int npi, const double* ai_, const double* ci_, int lj, int npj,
                                    const double* aj_, const double* cj_, const double* vec, double intcut,
                                    double (&out)[9][10]) {
  const double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
  const int nsi = 2 * li + 1, nsj = 2 * lj + 1;
#pragma unroll
  for (int b = 0; b < 9; b++)
#pragma unroll
    for (int x = 0; x < 10; x++) out[b][x] = 0.0;
  const double sqrtpi3 = 5.568327996831708;  // pi^(3/2)
  for (int ip = 0; ip < npi; ip++) {
    const double ai = ai_[ip];
    for (int jp = 0; jp < npj; jp++) {
      const double aj = aj_[jp];
      const double eab = ai + aj, oab = 1.0 / eab, est = ai * aj * r2 * oab;
      if (est > intcut) continue;
      const double so = sqrt(oab);
      const double pre = exp(-est) * sqrtpi3 * (so * so * so);
      const double cc = ci_[ip] * cj_[jp] * pre;
      double t[3][4][2];
#pragma unroll
      for (int k = 0; k < 3; k++) os_table<3, 1>(-vec[k] * aj * oab, vec[k] * ai * oab, 0.5 * oab, t[k]);
      for (int im = 0; im < nsi; im++) {
        const int ic = p_cart(li, im);
        for (int jm = 0; jm < nsj; jm++) {
          const int jc = p_cart(lj, jm);
          double v[3][3];
#pragma unroll
          for (int k = 0; k < 3; k++) {
            const int a = (ic == k), b = (jc == k);
#pragma unroll
            for (int m = 0; m < 3; m++) v[k][m] = t[k][a + m][b];
          }
          double* o = out[jm + nsj * im];
          o[0] += cc * v[0][0] * v[1][0] * v[2][0];
          o[1] += cc * v[0][1] * v[1][0] * v[2][0];
          o[2] += cc * v[0][0] * v[1][1] * v[2][0];
          o[3] += cc * v[0][0] * v[1][0] * v[2][1];
          o[4] += cc * v[0][2] * v[1][0] * v[2][0];
          o[5] += cc * v[0][1] * v[1][1] * v[2][0];
          o[6] += cc * v[0][0] * v[1][2] * v[2][0];
          o[7] += cc * v[0][1] * v[1][0] * v[2][1];
          o[8] += cc * v[0][0] * v[1][1] * v[2][1];
          o[9] += cc * v[0][0] * v[1][0] * v[2][2];
        }
      }
    }
  }
#pragma unroll
  for (int b = 0; b < 9; b++) traceless6(&out[b][4]);
}

// shellpair_multipole with the angular momenta fixed at compile time (blocks in registers).
template <int LI, int LJ>
__device__ __forceinline__ void shellpair_multipole_t(int npi, const double* ai_, const double* ci_, int npj,
                                                      const double* aj_, const double* cj_, const double* vec,
                                                      double intcut, double (&out)[(2 * LI + 1) * (2 * LJ + 1)][10]) {
  constexpr int NSI = 2 * LI + 1, NSJ = 2 * LJ + 1;
  const double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
#pragma unroll
  for (int b = 0; b < NSI * NSJ; b++)
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
      double t[3][4][2];
#pragma unroll
      for (int k = 0; k < 3; k++) os_table<3, 1>(-vec[k] * aj * oab, vec[k] * ai * oab, 0.5 * oab, t[k]);
#pragma unroll
      for (int im = 0; im < NSI; im++) {
        const int ic = LI == 0 ? -1 : (im == 0 ? 1 : (im == 1 ? 2 : 0));
#pragma unroll
        for (int jm = 0; jm < NSJ; jm++) {
          const int jc = LJ == 0 ? -1 : (jm == 0 ? 1 : (jm == 1 ? 2 : 0));
          double v[3][3];
#pragma unroll
          for (int k = 0; k < 3; k++) {
            const int a = (ic == k), b = (jc == k);
#pragma unroll
            for (int m = 0; m < 3; m++) v[k][m] = t[k][a + m][b];
          }
          double* o = out[jm + NSJ * im];
          o[0] += cc * v[0][0] * v[1][0] * v[2][0];
          o[1] += cc * v[0][1] * v[1][0] * v[2][0];
          o[2] += cc * v[0][0] * v[1][1] * v[2][0];
          o[3] += cc * v[0][0] * v[1][0] * v[2][1];
          o[4] += cc * v[0][2] * v[1][0] * v[2][0];
          o[5] += cc * v[0][1] * v[1][1] * v[2][0];
          o[6] += cc * v[0][0] * v[1][2] * v[2][0];
          o[7] += cc * v[0][1] * v[1][0] * v[2][1];
          o[8] += cc * v[0][0] * v[1][1] * v[2][1];
          o[9] += cc * v[0][0] * v[1][0] * v[2][2];
        }
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NSI * NSJ; b++) traceless6(&out[b][4]);
}

// Per-AO-pair coefficients of the gradient contraction of one shell pair (hamiltonian.rs:169-211):
// dE/dR_i = sum_blk [ sval ds + 2 hp s dsv - p (ddi.vdp_i + ddj.vdp_j + dqi.vqp_i + dqj.vqp_j) ]
struct GradCoef {
  double sval[9], hp2[9], p[9];
};

// Contract the ket-centre derivatives of S, D and Q of a shell pair (l <= 1) with `cf` and the
// multipole potentials of the two atoms; adds to dg[3] and returns sum_blk p s (for dE/dCN).
// The derivatives are those of integrals.rs:400-521, contracted per primitive pair (all terms are
// linear in the primitive contributions).
__device__ double shellpair_grad(int li, int npi, const double* ai_, const double* ci_, int lj, int npj,
                                 const double* aj_, const double* cj_, const double* vec, double intcut,
                                 const GradCoef& cf, const double* dsv, const double* vdi, const double* vdj,
                                 const double* vqi, const double* vqj, double* dg) {
  const double r2 = vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2];
  const int nsi = 2 * li + 1, nsj = 2 * lj + 1;
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
      const double pre = exp(-est) * sqrtpi3 * (so * so * so);
      const double cc = ci_[ip] * cj_[jp] * pre;
      double t[3][3][4];
#pragma unroll
      for (int k = 0; k < 3; k++) os_table<2, 3>(-vec[k] * aj * oab, vec[k] * ai * oab, 0.5 * oab, t[k]);
      for (int im = 0; im < nsi; im++) {
        const int ic = p_cart(li, im);
        for (int jm = 0; jm < nsj; jm++) {
          const int jc = p_cart(lj, jm);
          const int blk = jm + nsj * im;
          // moments around R_j and their R_i derivatives, per direction
          double v[3][3], g[3][3];
#pragma unroll
          for (int k = 0; k < 3; k++) {
            const int a = (ic == k), b = (jc == k);
#pragma unroll
            for (int m = 0; m < 3; m++) {
              v[k][m] = t[k][a][b + m];
              g[k][m] = ai2 * t[k][a + 1][b + m] - (a > 0 ? t[k][0][b + m] : 0.0);
            }
          }
          // value/derivative of s, dj (3), mj (6 raw second moments xx xy yy xz yz zz)
          const int MOM[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0},
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
          const double pij = cf.p[blk];
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
            const int AB[6][2] = {{0, 0}, {0, 1}, {1, 1}, {0, 2}, {1, 2}, {2, 2}};
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
            dg[k] += cf.sval[blk] * ds + cf.hp2[blk] * s * dsv[k] - pij * mp;
          }
        }
      }
    }
  }
  return ps;
}

// shellpair_grad with the angular momenta fixed at compile time (blocks in registers).
template <int LI, int LJ>
__device__ __forceinline__ double shellpair_grad_t(int npi, const double* ai_, const double* ci_, int npj,
                                                   const double* aj_, const double* cj_, const double* vec,
                                                   double intcut, const double* sval, const double* hp2,
                                                   const double* pp, const double* dsv, const double* vdi,
                                                   const double* vdj, const double* vqi, const double* vqj,
                                                   double* dg) {
  constexpr int NSI = 2 * LI + 1, NSJ = 2 * LJ + 1;
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
      double t[3][3][4];
#pragma unroll
      for (int k = 0; k < 3; k++) os_table<2, 3>(-vec[k] * aj * oab, vec[k] * ai * oab, 0.5 * oab, t[k]);
#pragma unroll
      for (int im = 0; im < NSI; im++) {
        const int ic = LI == 0 ? -1 : (im == 0 ? 1 : (im == 1 ? 2 : 0));
#pragma unroll
        for (int jm = 0; jm < NSJ; jm++) {
          const int jc = LJ == 0 ? -1 : (jm == 0 ? 1 : (jm == 1 ? 2 : 0));
          const int blk = jm + NSJ * im;
          double v[3][3], g[3][3];
#pragma unroll
          for (int k = 0; k < 3; k++) {
            const int a = (ic == k), b = (jc == k);
#pragma unroll
            for (int m = 0; m < 3; m++) {
              v[k][m] = t[k][a][b + m];
              g[k][m] = ai2 * t[k][a + 1][b + m] - (a > 0 ? t[k][0][b + m] : 0.0);
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
          const double pij = pp[blk];
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
            dg[k] += sval[blk] * ds + hp2[blk] * s * dsv[k] - pij * mp;
          }
        }
      }
    }
  }
  return ps;
}

}  // namespace axtb
