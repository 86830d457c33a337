// GFN2-xTB FP64 CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// A restatement of axtb's CPU path for GFN2-xTB energy + analytic gradient in axtb's default
// conventions (crates/axtb/src/conventions.rs:244-268): setup, integrals, SCC loop with Broyden
// mixing, Fermi smearing, energy partition and the full analytic gradient. It exists to check the
// HIP engine (tests, smoke(), bench.py's cpu_baseline leg) and is never linked into the product.
// Every function names the reference file:line it follows. Parameter tables come from
// gfn2_tables.h (generated data, see gen_tables.py).
//
// Pinned by: axtb's committed energies and per-term energies (tests/reference_matrix/results.json
// struct-water/-caffeine/-taxol, bench/results/final-9985wx/gfn2-eg-t1.json), the D4 fixtures
// (crates/d4/tests/fixtures/*.txt: CN, EEQ charges, C6, polarisabilities, 2-/3-body energies),
// central finite differences of the energy (gradient) and the invariants (sum g = 0,
// sum q = charge, Tr(PS) = N_el, self-overlap = 1). The eigensolver (Householder + implicit QL)
// is pinned by residual/orthonormality checks; faer's bits are not reproduced (parity unpinned at
// the last-bit level, which the 1e-8 Eh bar does not need).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include <omp.h>

#include "../axtb_amd/csrc/gfn2_tables.h"
#include "../axtb_amd/csrc/solvation.hpp"

namespace {

using std::vector;
typedef vector<double> vec;

// crates/base/src/units.rs:18-24 (CODATA 2022)
const double AUTOAA = 5.29177210544e-11 * 1e10;
const double AATOAU = 1.0 / AUTOAA;
const double KB_HARTREE = 3.1668115634564e-6;
const double KB_XTB_TBLITE = 3.166808578545117e-06;  // base/src/units.rs:81
const double PI = 3.14159265358979323846;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ------------------------------------------------------------------ basis (basis.rs:20-81)
struct Cgto {
  int l = 0, nprim = 0;
  double alpha[6] = {0}, coeff[6] = {0};
  double min_alpha() const {
    double a = INFINITY;
    for (int i = 0; i < nprim; i++) a = std::min(a, alpha[i]);
    return a;
  }
};

inline int nsph(int l) { return 2 * l + 1; }
inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// crates/base/src/math.rs:108-111 (tblite association of (2/pi) alpha)
double primitive_norm_tblite(int l, double alpha) {
  static const double DFACT[5] = {1.0, 1.0, 3.0, 15.0, 105.0};
  double top = 2.0 / PI;
  return std::pow(top * alpha, 0.75) * std::pow(std::sqrt(4.0 * alpha), l) / std::sqrt(DFACT[l]);
}

// basis.rs:20-81 (STO-NG of Stewart, GFN2 tables)
Cgto slater_to_gauss(int ng, int n, int l, double zeta) {
  if ((n > 5 || n <= l) && !(n == 6 && ng == 6 && l <= 1)) throw Error(1, "STO-NG: unsupported n,l");
  int ityp;
  switch (l) {
    case 0: ityp = n; break;
    case 1: ityp = 4 + n; break;
    case 2: ityp = 7 + n; break;
    default: throw Error(1, "STO-NG: l > 2 not supported here");
  }
  ityp -= 1;
  Cgto c;
  c.l = l;
  c.nprim = ng;
  double z2 = zeta * zeta;
  for (int i = 0; i < ng; i++) {
    double a, k;
    switch (ng) {
      case 1: a = STO_P_ALPHA1[ityp]; k = 1.0; break;
      case 2: a = STO_P_ALPHA2[ityp][i]; k = STO_P_COEFF2[ityp][i]; break;
      case 3: a = STO_P_ALPHA3[ityp][i]; k = STO_P_COEFF3[ityp][i]; break;
      case 4: a = STO_P_ALPHA4[ityp][i]; k = STO_P_COEFF4[ityp][i]; break;
      case 5: a = STO_P_ALPHA5[ityp][i]; k = STO_P_COEFF5[ityp][i]; break;
      case 6:  // n = 6: the dedicated 6s / 6p expansions (basis.rs:56-61)
        a = n == 6 ? (l == 0 ? STO_P_ALPHA_6S[i] : STO_P_ALPHA_6P[i]) : STO_P_ALPHA6[ityp][i];
        k = n == 6 ? (l == 0 ? STO_P_COEFF_6S[i] : STO_P_COEFF_6P[i]) : STO_P_COEFF6[ityp][i];
        break;
      default: throw Error(1, "STO-NG: bad primitive count");
    }
    c.alpha[i] = a * z2;
    c.coeff[i] = k;
  }
  for (int i = 0; i < ng; i++) c.coeff[i] *= primitive_norm_tblite(l, c.alpha[i]);
  return c;
}

// element tables of the axtb conventions (param/mod.rs:39-47, conventions.rs:244-268)
double pauling_en(int z) { return EL_PAULING_EN[z - 1]; }
double covalent_rad_d3_codata2022(int z) { return 4.0 / 3.0 * (EL_COVALENT_RAD_2009_AA[z - 1] * AATOAU); }
double atomic_rad_codata2022(int z) { return EL_ATOMIC_RAD_AA[z - 1] * AATOAU; }
// element_data.rs:94-102: mctc bohr, 4/3 of the 2009 radii (the EEQ guess's CN radii)
double covalent_rad_d3_mctc(int z) { return 4.0 / 3.0 * (AXTB_MCTC_AATOAU * EL_COVALENT_RAD_2009_AA[z - 1]); }

inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// ------------------------------------------------------------------ molecule + model
struct Model {
  int nat = 0, nsh = 0, nao = 0;
  vector<int> z;                 // atomic numbers
  vec xyz;                       // bohr, [nat][3]
  double charge = 0;
  // basis maps (basis.rs:165-229)
  vector<vector<Cgto>> shells;   // per atom
  vector<int> ish_at, nsh_at, iao_at, nao_at, sh2at, iao_sh, nao_sh, ao2at, ao2sh;
  vec min_alpha_sh, min_alpha_at;
  double intcut = 0;
  // conventions (conventions.rs:244-353): tblite 0.8.0's (mctc-lib radii, 25 bohr CN and repulsion
  // cutoffs, atom-pair screening of the integrals) or axtb's
  bool tblite = false;
  vec rcov;                      // GFN2 CN radius per atom
  double pair_cut2 = INFINITY, cn_cut2 = 1600.0, rep_cut2 = 1600.0;
  // H0 (gfn2.rs:61-99)
  vec se0, kcn, shpoly, refocc;  // per shell
  vec rad;                       // per atom
  vector<int> ang;               // per shell
  vec zeta_sh;                   // Slater exponent per shell
};

const double WEXP = 0.5, ENSCALE = 2.0e-2;
const double KDIAG[5] = {1.85, 2.23, 2.23, 2.23, 2.23};
const double SHELL_HUBBARD_DERIVS[5] = {1.0, 0.5, 0.25, 0.25, 0.25};

// gfn2.rs:24-30
double kshell(int k, int l) {
  if ((k == 2 && (l == 0 || l == 1)) || (l == 2 && (k == 0 || k == 1))) return 2.0;
  return (KDIAG[l] + KDIAG[k]) / 2.0;
}

// hscale of the shell pair (shell ish of atom i, jsh of atom j), gfn2.rs:76-98, used as
// hscale[jsp][isp][jsh][ish] by hamiltonian.rs:88
double hscale(const Model& m, int zi, int ish, int zj, int jsh) {
  double den = std::pow(pauling_en(zi) - pauling_en(zj), 2);
  int il = GFN2_ANG_SHELL[zi - 1][ish], jl = GFN2_ANG_SHELL[zj - 1][jsh];
  double za = GFN2_SLATER_EXPONENT[zi - 1][ish], zb = GFN2_SLATER_EXPONENT[zj - 1][jsh];
  double zij = std::pow(2.0 * std::sqrt(za * zb) / (za + zb), WEXP);
  double enp = 1.0 + ENSCALE * den;
  double km = kshell(jl, il) * enp;
  return zij * km;
}

Model build_model(int nat, const int* z, const double* xyz, double charge, double acc, bool tblite = false) {
  Model m;
  m.tblite = tblite;
  m.nat = nat;
  m.z.assign(z, z + nat);
  m.xyz.assign(xyz, xyz + 3 * nat);
  m.charge = charge;
  for (int i = 0; i < nat; i++)
    if (z[i] < 1 || z[i] > AXTB_GFN2_MAX_ELEM) throw Error(4, "unsupported element");
  // molecule validation (base/src/molecule.rs:12,35-64): atoms closer than 1e-3 bohr... keep the
  // distance check the reference does
  for (int i = 0; i < nat; i++)
    for (int j = 0; j < i; j++) {
      double d2 = 0;
      for (int k = 0; k < 3; k++) d2 += std::pow(xyz[3 * i + k] - xyz[3 * j + k], 2);
      if (!(d2 >= 1e-6)) throw Error(1, "atoms too close");
    }
  int ish = 0, iao = 0;
  for (int i = 0; i < nat; i++) {
    int iz = z[i] - 1;
    vector<Cgto> sh;
    for (int s = 0; s < GFN2_NSHELL[iz]; s++) {
      int l = GFN2_ANG_SHELL[iz][s];
      if (l > 2) throw Error(4, "f shells unsupported");
      sh.push_back(slater_to_gauss(GFN2_NUMBER_OF_PRIMITIVES[iz][s], GFN2_PRINCIPAL_QUANTUM_NUMBER[iz][s], l,
                                   GFN2_SLATER_EXPONENT[iz][s]));
    }
    m.ish_at.push_back(ish);
    m.nsh_at.push_back((int)sh.size());
    m.iao_at.push_back(iao);
    double amin = INFINITY;
    for (size_t s = 0; s < sh.size(); s++) {
      m.sh2at.push_back(i);
      m.iao_sh.push_back(iao);
      m.nao_sh.push_back(nsph(sh[s].l));
      double a = sh[s].min_alpha();
      m.min_alpha_sh.push_back(a);
      amin = std::min(amin, a);
      for (int k = 0; k < nsph(sh[s].l); k++) {
        m.ao2at.push_back(i);
        m.ao2sh.push_back(ish);
      }
      iao += nsph(sh[s].l);
      ish++;
      int l = sh[s].l;
      m.ang.push_back(l);
      m.se0.push_back(GFN2_P_SELFENERGY[iz][s]);
      m.kcn.push_back(GFN2_P_KCN[iz][l]);
      m.shpoly.push_back(GFN2_P_SHPOLY[iz][l]);
      m.refocc.push_back(GFN2_REFERENCE_OCC[iz][l]);
      m.zeta_sh.push_back(GFN2_SLATER_EXPONENT[iz][s]);
    }
    m.nao_at.push_back(iao - m.iao_at[i]);
    m.min_alpha_at.push_back(amin);
    m.shells.push_back(sh);
    // radii: CODATA 2022 bohr (axtb) or mctc-lib's (tblite: element_data.rs:26,94-102)
    m.rad.push_back(tblite ? AXTB_MCTC_AATOAU * EL_ATOMIC_RAD_AA[z[i] - 1] : atomic_rad_codata2022(z[i]));
    m.rcov.push_back(tblite ? covalent_rad_d3_mctc(z[i]) : covalent_rad_d3_codata2022(z[i]));
  }
  m.nsh = ish;
  m.nao = iao;
  if (tblite) {
    // tblite: primitive screening at accuracy 1, atom pairs within sqrt(2 intcut(acc) / alpha_min)
    // (at most 40 bohr), CN and repulsion within 25 bohr (basis.rs:139-143,207,247-249;
    // conventions.rs:338-346)
    auto integral_cutoff = [](double a) {
      a = std::min(std::max(a, 1.0e-4), 1.0e3);
      return std::min(std::max(25.0 - 10.0 * std::log10(a), 5.0), 25.0);
    };
    m.intcut = integral_cutoff(1.0);
    double amin = INFINITY;
    for (double a : m.min_alpha_sh) amin = std::min(amin, a);
    const double rc = std::min(std::sqrt(2.0 * integral_cutoff(acc) / amin), 40.0);
    m.pair_cut2 = rc * rc;
    m.cn_cut2 = m.rep_cut2 = 25.0 * 25.0;
  } else {
    // basis.rs:135-137 (xtb's cutoff, axtb conventions: primitive_cutoff = integral_cutoff_xtb)
    m.intcut = std::max(25.0 - 10.0 * std::log10(acc), 20.0);
  }
  return m;
}

// ------------------------------------------------------------------ coordination numbers
// base/src/ncoord.rs:68-95: double exponential counting function of GFN2
inline double exp_count(double k, double r, double r0) { return 1.0 / (1.0 + std::exp(-k * (r0 / r - 1.0))); }
inline double exp_dcount(double k, double r, double r0) {
  double e = std::exp(-k * (r0 / r - 1.0));
  return (-k * r0 * e) / (r * r * (e + 1.0) * (e + 1.0));
}
inline double dexp_count(double r, double rc) { return exp_count(10.0, r, rc) * exp_count(20.0, r, rc + 2.0); }
inline double dexp_dcount(double r, double rc) {
  return exp_dcount(10.0, r, rc) * exp_count(20.0, r, rc + 2.0) + exp_count(10.0, r, rc) * exp_dcount(20.0, r, rc + 2.0);
}
// base/src/ncoord.rs:140-152 (dftd4 erf_count, k = 7.5)
const double SQRTPI = 1.7724538509055159;
inline double erf_count(double k, double r, double rc) { return 0.5 * (1.0 + std::erf(-((k * (r - rc)) / rc))); }
inline double erf_dcount(double k, double r, double rc) {
  double e = k * (r - rc) / rc;
  return -(k * std::exp(-(e * e))) / (SQRTPI * rc);
}

// GFN2 CN: axtb/src/ncoord.rs:29-41 over base/src/ncoord.rs:471-484 (cutoff 40, tblite 25; r2 < 1e-12
// skipped)
vec gfn2_cn(const Model& m) {
  vec cn(m.nat, 0.0);
  double c2 = m.cn_cut2;
  for (int i = 0; i < m.nat; i++)
    for (int j = 0; j < i; j++) {
      const double* a = &m.xyz[3 * i];
      const double* b = &m.xyz[3 * j];
      double v[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
      double r2 = dot3(v, v);
      if (r2 > c2 || r2 < 1e-12) continue;
      double rc = m.rcov[i] + m.rcov[j];
      double c = dexp_count(std::sqrt(r2), rc);
      cn[i] += c;
      cn[j] += c;
    }
  return cn;
}

// base/src/ncoord.rs:486-519 (GradientOrder::Scaled)
void gfn2_cn_gradient(const Model& m, const vec& dedcn, vec& g) {
  double c2 = m.cn_cut2;
  for (int i = 0; i < m.nat; i++)
    for (int j = 0; j < i; j++) {
      double s = dedcn[i] + dedcn[j];
      const double* a = &m.xyz[3 * i];
      const double* b = &m.xyz[3 * j];
      double v[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
      double r2 = dot3(v, v);
      if (r2 > c2 || r2 < 1e-12) continue;
      double r = std::sqrt(r2);
      double rc = m.rcov[i] + m.rcov[j];
      double t = dexp_dcount(r, rc) * s / r;
      for (int k = 0; k < 3; k++) {
        g[3 * i + k] += t * v[k];
        g[3 * j + k] -= t * v[k];
      }
    }
}

// ------------------------------------------------------------------ dense helpers
// column-major n x n
struct Mat {
  int n = 0, m = 0;
  vec a;
  Mat() {}
  Mat(int n_, int m_) : n(n_), m(m_), a((size_t)n_ * m_, 0.0) {}
  double& operator()(int i, int j) { return a[(size_t)j * n + i]; }
  double operator()(int i, int j) const { return a[(size_t)j * n + i]; }
};

// C = alpha op(A) op(B) + beta C, column-major (op = transpose when ta / tb). Tiles of C over the
// OpenMP threads; each tile packs its panels of A and B and sums K in ascending chunks, so the
// result does not depend on the thread count. Plain blocked loops (the reference multiplies with
// faer's blocked products; the bits of either are not pinned).
void mm(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb,
        double beta, double* C, int ldc) {
  if (m <= 0 || n <= 0) return;
  const int MC = 96, NC = 96, KC = 256, MR = 8, NR = 4;
  const int tm = (m + MC - 1) / MC, tn = (n + NC - 1) / NC;
#pragma omp parallel
  {
    vector<double> Ap((size_t)MC * KC), Bp((size_t)KC * NC);
#pragma omp for collapse(2) schedule(static)
    for (int bi = 0; bi < tm; bi++)
      for (int bj = 0; bj < tn; bj++) {
        const int i0 = bi * MC, j0 = bj * NC, mc = std::min(MC, m - i0), nc = std::min(NC, n - j0);
        for (int j = 0; j < nc; j++) {
          double* c = C + (size_t)(j0 + j) * ldc + i0;
          for (int i = 0; i < mc; i++) c[i] = beta == 0.0 ? 0.0 : beta * c[i];
        }
        for (int p0 = 0; p0 < k; p0 += KC) {
          const int kc = std::min(KC, k - p0);
          for (int ir = 0; ir < mc; ir += MR)
            for (int p = 0; p < kc; p++)
              for (int i = 0; i < MR; i++) {
                const int gi = i0 + ir + i, gp = p0 + p;
                Ap[(size_t)ir * kc + p * MR + i] =
                    ir + i < mc ? (ta ? A[(size_t)gi * lda + gp] : A[(size_t)gp * lda + gi]) : 0.0;
              }
          for (int jr = 0; jr < nc; jr += NR)
            for (int p = 0; p < kc; p++)
              for (int j = 0; j < NR; j++) {
                const int gj = j0 + jr + j, gp = p0 + p;
                Bp[(size_t)jr * kc + p * NR + j] =
                    jr + j < nc ? (tb ? B[(size_t)gp * ldb + gj] : B[(size_t)gj * ldb + gp]) : 0.0;
              }
          for (int jr = 0; jr < nc; jr += NR)
            for (int ir = 0; ir < mc; ir += MR) {
              double acc[NR][MR] = {};
              const double* a = &Ap[(size_t)ir * kc];
              const double* b = &Bp[(size_t)jr * kc];
              for (int p = 0; p < kc; p++)
                for (int j = 0; j < NR; j++)
                  for (int i = 0; i < MR; i++) acc[j][i] += a[p * MR + i] * b[p * NR + j];
              for (int j = 0; j < NR && jr + j < nc; j++) {
                double* c = C + (size_t)(j0 + jr + j) * ldc + i0 + ir;
                for (int i = 0; i < MR && ir + i < mc; i++) c[i] += alpha * acc[j][i];
              }
            }
        }
      }
  }
}

// the lower triangle of a column-major n x n matrix copied to the upper one
void mirror_lower(Mat& M) {
  const int n = M.n;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (int i = 0; i < j; i++) M(i, j) = M(j, i);
}

// Gaussian elimination with partial pivoting on the last maximum (base/src/linalg/lu.rs:81-107)
bool gauss_last_max(vector<vec>& a, vec& b) {
  int n = (int)b.size();
  for (int col = 0; col < n; col++) {
    int piv = col;
    for (int i = col; i < n; i++)
      if (std::fabs(a[i][col]) >= std::fabs(a[piv][col])) piv = i;
    if (a[piv][col] == 0.0) return false;
    std::swap(a[col], a[piv]);
    std::swap(b[col], b[piv]);
    for (int row = col + 1; row < n; row++) {
      double f = a[row][col] / a[col][col];
      if (f != 0.0) {
        for (int k = col; k < n; k++) a[row][k] -= f * a[col][k];
        b[row] -= f * b[col];
      }
    }
  }
  for (int col = n - 1; col >= 0; col--) {
    double s = b[col];
    for (int k = col + 1; k < n; k++) s -= a[col][k] * b[k];
    b[col] = s / a[col][col];
  }
  return true;
}

// dense LU with partial pivoting (first maximum) for the EEQ system (faer's Lu in base/eeq.rs)
vec lu_solve(Mat A, vec b) {
  int n = A.n;
  vector<int> perm(n);
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int i = k + 1; i < n; i++)
      if (std::fabs(A(i, k)) > std::fabs(A(p, k))) p = i;
    if (A(p, k) == 0.0) throw Error(2, "singular EEQ system");
    if (p != k) {
      for (int j = 0; j < n; j++) std::swap(A(k, j), A(p, j));
      std::swap(b[k], b[p]);
    }
    double piv = A(k, k);
    for (int i = k + 1; i < n; i++) A(i, k) /= piv;
#pragma omp parallel for schedule(static, 8) if (n - k > 256)
    for (int j = k + 1; j < n; j++) {
      double akj = A(k, j);
      if (akj == 0.0) continue;
      double* col = &A.a[(size_t)j * n];
      const double* lk = &A.a[(size_t)k * n];
      for (int i = k + 1; i < n; i++) col[i] -= lk[i] * akj;
    }
  }
  for (int i = 0; i < n; i++) {
    double s = b[i];
    for (int k = 0; k < i; k++) s -= A(i, k) * b[k];
    b[i] = s;
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = b[i];
    for (int k = i + 1; k < n; k++) s -= A(i, k) * b[k];
    b[i] = s / A(i, i);
  }
  return b;
}

// ------------------------------------------------------------------ EEQ guess (axtb/src/eeq.rs:43-77)
inline double log_cap(double cn, double cnmax) {  // base/src/ncoord.rs:258-262
  double a = std::log(1.0 + std::exp(cnmax));
  return a - std::log(std::exp(cnmax - cn) + 1.0);
}

vec eeq_charges(int nat, const int* z, const double* xyz, double charge) {
  for (int i = 0; i < nat; i++)
    if (z[i] > 103) throw Error(4, "EEQ: unsupported element");
  vec cn(nat, 0.0);
  double c2 = 25.0 * 25.0;
  for (int i = 0; i < nat; i++)
    for (int j = 0; j < i; j++) {
      double v[3] = {xyz[3 * i] - xyz[3 * j], xyz[3 * i + 1] - xyz[3 * j + 1], xyz[3 * i + 2] - xyz[3 * j + 2]};
      double r2 = dot3(v, v);
      if (r2 > c2 || r2 < 1e-12) continue;
      double rc = covalent_rad_d3_mctc(z[i]) + covalent_rad_d3_mctc(z[j]);
      double c = erf_count(7.5, std::sqrt(r2), rc);
      cn[i] += c;
      cn[j] += c;
    }
  for (int i = 0; i < nat; i++) cn[i] = log_cap(cn[i], 8.0);
  double sqrt2pi = std::sqrt(2.0 / PI);
  Mat A(nat + 1, nat + 1);
  for (int i = 0; i < nat; i++) {
    int zi = z[i] - 1;
    A(i, i) = EEQ_ETA[zi] + sqrt2pi / EEQ_RAD[zi];
    for (int j = 0; j < i; j++) {
      int zj = z[j] - 1;
      double v[3] = {xyz[3 * i] - xyz[3 * j], xyz[3 * i + 1] - xyz[3 * j + 1], xyz[3 * i + 2] - xyz[3 * j + 2]};
      double r2 = dot3(v, v);
      double gam = 1.0 / (EEQ_RAD[zi] * EEQ_RAD[zi] + EEQ_RAD[zj] * EEQ_RAD[zj]);
      double val = std::erf(std::sqrt(r2 * gam)) / std::sqrt(r2);
      A(i, j) = val;
      A(j, i) = val;
    }
    A(i, nat) = 1.0;
    A(nat, i) = 1.0;
  }
  vec x(nat + 1);
  for (int i = 0; i < nat; i++) {
    int zi = z[i] - 1;
    x[i] = -EEQ_CHI[zi] + EEQ_KCNCHI[zi] / std::sqrt(cn[i] + 1e-14) * cn[i];
  }
  x[nat] = charge;
  vec q = lu_solve(A, x);
  q.resize(nat);
  for (double v : q)
    if (!std::isfinite(v)) throw Error(2, "EEQ charges are not finite");
  return q;
}

// ------------------------------------------------------------------ one-electron integrals
// integrals.rs:29-45: Cartesian exponents in CCA order
void cart_exps(int l, int out[][3]) {
  int c = 0;
  for (int lx = l; lx >= 0; lx--)
    for (int ly = l - lx; ly >= 0; ly--) {
      out[c][0] = lx;
      out[c][1] = ly;
      out[c][2] = l - lx - ly;
      c++;
    }
}

// integrals.rs:49-103 (s, p, d rows): spherical rows from Cartesian columns
void sph_trafo(int l, double T[5][6]) {
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 6; j++) T[i][j] = 0.0;
  if (l == 0) {
    T[0][0] = 1.0;
  } else if (l == 1) {  // p: y, z, x from x, y, z
    T[0][1] = 1.0;
    T[1][2] = 1.0;
    T[2][0] = 1.0;
  } else {
    double s3 = std::sqrt(3.0), s3_4 = s3 * 0.5;
    T[0][1] = s3;
    T[1][4] = s3;
    T[2][0] = -0.5; T[2][3] = -0.5; T[2][5] = 1.0;
    T[3][2] = s3;
    T[4][0] = s3_4; T[4][3] = -s3_4;
  }
}

const int MT = 6;  // 1D table size (l <= 2, +2 for moments, +1 for derivatives)
// 1D moment table t[i][j] = <(x-A)^i (x-B)^j> / (Gaussian prefactor), Obara-Saika recursions
// (integrals.rs:103-127): rpa = P - A, rpb = P - B, h = 1/(2(a+b))
void table_1d(int la, int lb, double rpa, double rpb, double oab, double t[MT][MT]) {
  double h = 0.5 * oab;
  for (int i = 0; i < MT; i++)
    for (int j = 0; j < MT; j++) t[i][j] = 0.0;
  t[0][0] = 1.0;
  for (int i = 0; i < la; i++) t[i + 1][0] = rpa * t[i][0] + (i > 0 ? i * h * t[i - 1][0] : 0.0);
  for (int j = 0; j < lb; j++)
    for (int i = 0; i <= la; i++) {
      double v = rpb * t[i][j];
      if (i > 0) v += i * h * t[i - 1][j];
      if (j > 0) v += j * h * t[i][j - 1];
      t[i][j + 1] = v;
    }
}

inline bool pair_screened(double a, double b, double r2, double intcut) {  // integrals.rs:216-218
  return a * b * r2 / (a + b) > intcut * (1.0 + 1e-10);
}

// traceless form Q = 3/2 M - 1/2 tr(M), packed xx xy yy xz yz zz (integrals.rs:233-241)
inline void traceless(double* q) {
  double tr = 0.5 * (q[0] + q[2] + q[5]);
  q[0] = 1.5 * q[0] - tr;
  q[1] *= 1.5;
  q[2] = 1.5 * q[2] - tr;
  q[3] *= 1.5;
  q[4] *= 1.5;
  q[5] = 1.5 * q[5] - tr;
}

// moment exponents of s, x, y, z, xx, xy, yy, xz, yz, zz
const int MOM[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0},
                        {1, 1, 0}, {0, 2, 0}, {1, 0, 1}, {0, 1, 1}, {0, 0, 2}};

// Cartesian -> spherical of NV interleaved components: out[(jm + nsj*im)*NV + x]
void to_sph(int lj, int li, int NV, const double* cart, double* out) {
  double Tj[5][6], Ti[5][6];
  sph_trafo(lj, Tj);
  sph_trafo(li, Ti);
  int ncj = ncart(lj), nci = ncart(li), nsj = nsph(lj), nsi = nsph(li);
  for (int im = 0; im < nsi; im++)
    for (int jm = 0; jm < nsj; jm++)
      for (int x = 0; x < NV; x++) {
        double acc = 0.0;
        for (int ic = 0; ic < nci; ic++) {
          if (Ti[im][ic] == 0.0) continue;
          for (int jc = 0; jc < ncj; jc++) {
            if (Tj[jm][jc] == 0.0) continue;
            acc += Ti[im][ic] * Tj[jm][jc] * cart[(jc + ncj * ic) * NV + x];
          }
        }
        out[(jm + nsj * im) * NV + x] = acc;
      }
}

// Overlap, dipole and traceless quadrupole of a shell pair, moments around the ket centre R_i
// (integrals.rs:277-367). Bra cj on R_j, ket ci on R_i, vec = R_i - R_j. Blocks [jm + nsj*im].
void multipole(const Cgto& cj, const Cgto& ci, const double* vec, double intcut, double* s, double (*d)[3],
               double (*q)[6]) {
  double r2 = dot3(vec, vec);
  int li = ci.l, lj = cj.l;
  int ei[6][3], ej[6][3];
  cart_exps(li, ei);
  cart_exps(lj, ej);
  int ncj = ncart(lj), nci = ncart(li);
  double cart[36 * 10];
  std::fill(cart, cart + 360, 0.0);
  double t[3][MT][MT];
  for (int ip = 0; ip < ci.nprim; ip++)
    for (int jp = 0; jp < cj.nprim; jp++) {
      double ai = ci.alpha[ip], aj = cj.alpha[jp];
      double eab = ai + aj, oab = 1.0 / eab, est = ai * aj * r2 * oab;
      if (est > intcut) continue;
      double pre = std::exp(-est) * std::pow(std::sqrt(PI), 3) * std::pow(std::sqrt(oab), 3);
      for (int k = 0; k < 3; k++) table_1d(li + 2, lj, -vec[k] * aj * oab, vec[k] * ai * oab, oab, t[k]);
      double cc = ci.coeff[ip] * cj.coeff[jp] * pre;
      for (int ic = 0; ic < nci; ic++)
        for (int jc = 0; jc < ncj; jc++) {
          double* c = &cart[(jc + ncj * ic) * 10];
          for (int n = 0; n < 10; n++)
            c[n] += cc * t[0][ei[ic][0] + MOM[n][0]][ej[jc][0]] * t[1][ei[ic][1] + MOM[n][1]][ej[jc][1]] *
                    t[2][ei[ic][2] + MOM[n][2]][ej[jc][2]];
        }
    }
  double sph[25 * 10];
  to_sph(lj, li, 10, cart, sph);
  for (int k = 0; k < nsph(lj) * nsph(li); k++) {
    const double* c = &sph[k * 10];
    s[k] = c[0];
    for (int x = 0; x < 3; x++) d[k][x] = c[1 + x];
    for (int x = 0; x < 6; x++) q[k][x] = c[4 + x];
    traceless(q[k]);
  }
}

// Derivatives with respect to the ket centre R_i (integrals.rs:400-521)
struct MpGrad {
  double s[25], ds[25][3], ddi[25][3][3], ddj[25][3][3], dqi[25][3][6], dqj[25][3][6];
};

void multipole_grad(const Cgto& cj, const Cgto& ci, const double* vec, double intcut, MpGrad& out) {
  double r2 = dot3(vec, vec);
  int li = ci.l, lj = cj.l;
  int ei[6][3], ej[6][3];
  cart_exps(li, ei);
  cart_exps(lj, ej);
  int ncj = ncart(lj), nci = ncart(li);
  // per Cartesian pair: values [10] (moments around R_j) and derivatives [10][3]
  double val[36 * 10], der[36 * 30];
  std::fill(val, val + 360, 0.0);
  std::fill(der, der + 1080, 0.0);
  double t[3][MT][MT];
  for (int ip = 0; ip < ci.nprim; ip++)
    for (int jp = 0; jp < cj.nprim; jp++) {
      double ai = ci.alpha[ip], aj = cj.alpha[jp];
      double eab = ai + aj, oab = 1.0 / eab, est = ai * aj * r2 * oab;
      if (est > intcut) continue;
      double pre = std::exp(-est) * std::pow(std::sqrt(PI), 3) * std::pow(std::sqrt(oab), 3);
      for (int k = 0; k < 3; k++) table_1d(li + 1, lj + 2, -vec[k] * aj * oab, vec[k] * ai * oab, oab, t[k]);
      double cc = ci.coeff[ip] * cj.coeff[jp] * pre;
      double ai2 = 2.0 * ai;
      for (int ic = 0; ic < nci; ic++)
        for (int jc = 0; jc < ncj; jc++) {
          double v[3][3], g[3][3];
          for (int k = 0; k < 3; k++) {
            int a = ei[ic][k], b = ej[jc][k];
            for (int mm = 0; mm < 3; mm++) {
              v[k][mm] = t[k][a][b + mm];
              g[k][mm] = ai2 * t[k][a + 1][b + mm] - (a > 0 ? a * t[k][a - 1][b + mm] : 0.0);
            }
          }
          double* vc = &val[(jc + ncj * ic) * 10];
          double* dc = &der[(jc + ncj * ic) * 30];
          for (int n = 0; n < 10; n++) {
            double x = v[0][MOM[n][0]], y = v[1][MOM[n][1]], zz = v[2][MOM[n][2]];
            vc[n] += cc * x * y * zz;
            dc[3 * n] += cc * g[0][MOM[n][0]] * y * zz;
            dc[3 * n + 1] += cc * x * g[1][MOM[n][1]] * zz;
            dc[3 * n + 2] += cc * x * y * g[2][MOM[n][2]];
          }
        }
    }
  double sv[25 * 10], sd[25 * 30];
  to_sph(lj, li, 10, val, sv);
  to_sph(lj, li, 30, der, sd);
  static const int AB[6][2] = {{0, 0}, {0, 1}, {1, 1}, {0, 2}, {1, 2}, {2, 2}};
  for (int blk = 0; blk < nsph(lj) * nsph(li); blk++) {
    const double* v = &sv[blk * 10];
    const double* dv = &sd[blk * 30];
    double s = v[0];
    double dj[3] = {v[1], v[2], v[3]};
    double ds[3] = {dv[0], dv[1], dv[2]};
    out.s[blk] = s;
    for (int k = 0; k < 3; k++) out.ds[blk][k] = ds[k];
    double ddj[3][3], dmj[3][6], ddi[3][3], dmi[3][6];
    for (int k = 0; k < 3; k++) {
      for (int c = 0; c < 3; c++) ddj[k][c] = dv[3 * (1 + c) + k];
      for (int c = 0; c < 6; c++) dmj[k][c] = dv[3 * (4 + c) + k];
    }
    // shift to R_i: D_i = D_j - vec S; M_i = M_j - v_a D_j,b - v_b D_j,a + v_a v_b S
    for (int k = 0; k < 3; k++) {
      for (int c = 0; c < 3; c++) ddi[k][c] = ddj[k][c] - vec[c] * ds[k];
      ddi[k][k] -= s;
      for (int n = 0; n < 6; n++) {
        int a = AB[n][0], b = AB[n][1];
        double x = dmj[k][n] - vec[a] * ddj[k][b] - vec[b] * ddj[k][a] + vec[a] * vec[b] * ds[k];
        if (a == k) x += -dj[b] + vec[b] * s;
        if (b == k) x += -dj[a] + vec[a] * s;
        dmi[k][n] = x;
      }
    }
    for (int k = 0; k < 3; k++) {
      traceless(dmi[k]);
      traceless(dmj[k]);
    }
    std::memcpy(out.ddi[blk], ddi, sizeof ddi);
    std::memcpy(out.ddj[blk], ddj, sizeof ddj);
    std::memcpy(out.dqi[blk], dmi, sizeof dmi);
    std::memcpy(out.dqj[blk], dmj, sizeof dmj);
  }
}

// integrals.rs:991-1006: moments around the ket centre -> around the bra centre
void shift_to_bra(const double* vec, double s, const double* di, const double* qi, double* dj, double* qj) {
  for (int k = 0; k < 3; k++) dj[k] = di[k] + vec[k] * s;
  qj[0] = 2.0 * vec[0] * di[0] + vec[0] * vec[0] * s;
  qj[1] = vec[0] * di[1] + vec[1] * di[0] + vec[0] * vec[1] * s;
  qj[2] = 2.0 * vec[1] * di[1] + vec[1] * vec[1] * s;
  qj[3] = vec[0] * di[2] + vec[2] * di[0] + vec[0] * vec[2] * s;
  qj[4] = vec[1] * di[2] + vec[2] * di[1] + vec[1] * vec[2] * s;
  qj[5] = 2.0 * vec[2] * di[2] + vec[2] * vec[2] * s;
  traceless(qj);
  for (int k = 0; k < 6; k++) qj[k] += qi[k];
}

// ------------------------------------------------------------------ H0, S, D, Q (ao.rs:195-291, hamiltonian.rs:77-94)
struct Ints {
  Mat S, H0, D[3], Q[6];
};

Ints build_integrals(const Model& m, const vec& se) {
  int n = m.nao;
  Ints I;
  I.S = Mat(n, n);
  I.H0 = Mat(n, n);
  for (int k = 0; k < 3; k++) I.D[k] = Mat(n, n);
  for (int k = 0; k < 6; k++) I.Q[k] = Mat(n, n);
  // atom pairs write disjoint blocks (and their mirrors), so the rows of atoms run in parallel
#pragma omp parallel for schedule(dynamic, 4)
  for (int iat = 0; iat < m.nat; iat++) {
    double s[25], d[25][3], q[25][6];
    for (int jat = 0; jat <= iat; jat++) {
      double vec[3];
      for (int k = 0; k < 3; k++) vec[k] = m.xyz[3 * iat + k] - m.xyz[3 * jat + k];
      double r2 = dot3(vec, vec);
      bool onsite = iat == jat;
      if (r2 > m.pair_cut2) continue;  // the conventions' atom-pair cutoff (ao.rs:96-103)
      if (!onsite && pair_screened(m.min_alpha_at[iat], m.min_alpha_at[jat], r2, m.intcut)) continue;
      double rr = onsite ? 0.0 : std::sqrt(std::sqrt(r2) / (m.rad[jat] + m.rad[iat]));
      for (int ish = 0; ish < m.nsh_at[iat]; ish++) {
        int si = m.ish_at[iat] + ish;
        const Cgto& ci = m.shells[iat][ish];
        for (int jsh = 0; jsh < m.nsh_at[jat]; jsh++) {
          int sj = m.ish_at[jat] + jsh;
          const Cgto& cj = m.shells[jat][jsh];
          if (pair_screened(m.min_alpha_sh[si], m.min_alpha_sh[sj], r2, m.intcut)) continue;
          multipole(cj, ci, vec, m.intcut, s, d, q);
          double hij;
          if (onsite) {
            hij = 0.5 * (se[si] + se[sj]);
          } else {
            double sp = (1.0 + m.shpoly[si] * rr) * (1.0 + m.shpoly[sj] * rr);
            hij = 0.5 * (se[si] + se[sj]) * hscale(m, m.z[jat], jsh, m.z[iat], ish) * sp;
          }
          int nj = nsph(cj.l);
          for (int iao = 0; iao < nsph(ci.l); iao++)
            for (int jao = 0; jao < nj; jao++) {
              int ij = jao + nj * iao;
              int mu = m.iao_sh[sj] + jao, nu = m.iao_sh[si] + iao;
              I.S(mu, nu) = s[ij];
              I.H0(mu, nu) = s[ij] * hij;
              for (int k = 0; k < 3; k++) I.D[k](mu, nu) = d[ij][k];
              for (int k = 0; k < 6; k++) I.Q[k](mu, nu) = q[ij][k];
              if (!onsite) {
                double dj[3], qj[6];
                shift_to_bra(vec, s[ij], d[ij], q[ij], dj, qj);
                I.S(nu, mu) = s[ij];
                I.H0(nu, mu) = s[ij] * hij;
                for (int k = 0; k < 3; k++) I.D[k](nu, mu) = dj[k];
                for (int k = 0; k < 6; k++) I.Q[k](nu, mu) = qj[k];
              }
            }
        }
      }
    }
  }
  return I;
}

// ------------------------------------------------------------------ electrostatics (coulomb.rs)
struct Coulomb {
  vec hardness;     // per shell
  vec gamma3;       // per shell (ES3)
  vec gam;          // nsh x nsh row-major
  // AES (coulomb.rs:352-397)
  vec dkernel, qkernel, mrad, dmrdcn;  // per atom
};

const double QSCALE[6] = {1.0, 2.0, 1.0, 2.0, 2.0, 1.0};

Coulomb build_coulomb(const Model& m, const vec& cn) {
  Coulomb c;
  int nsh = m.nsh;
  c.hardness.resize(nsh);
  c.gamma3.resize(nsh);
  for (int s = 0; s < nsh; s++) {
    int iz = m.z[m.sh2at[s]] - 1;
    c.hardness[s] = GFN2_HUBBARD_PARAMETER[iz] * GFN2_SHELL_HUBBARD[iz][m.ang[s]];
    c.gamma3[s] = GFN2_P_HUBBARD_DERIVS[iz] * SHELL_HUBBARD_DERIVS[m.ang[s]];
  }
  // coulomb.rs:61-119 (gexp = 2)
  c.gam.assign((size_t)nsh * nsh, 0.0);
  for (int iat = 0; iat < m.nat; iat++)
    for (int jat = 0; jat <= iat; jat++) {
      double v[3];
      for (int k = 0; k < 3; k++) v[k] = m.xyz[3 * jat + k] - m.xyz[3 * iat + k];
      double r1g = std::pow(std::sqrt(dot3(v, v)), 2.0);
      for (int ish = 0; ish < m.nsh_at[iat]; ish++) {
        int jmax = iat == jat ? ish + 1 : m.nsh_at[jat];
        for (int jsh = 0; jsh < jmax; jsh++) {
          int a = m.ish_at[iat] + ish, b = m.ish_at[jat] + jsh;
          double eta = 0.5 * (c.hardness[a] + c.hardness[b]);
          double val = iat == jat ? eta : 1.0 / std::pow(r1g + std::pow(eta, -2.0), 0.5);
          c.gam[(size_t)a * nsh + b] = val;
          c.gam[(size_t)b * nsh + a] = val;
        }
      }
    }
  int nat = m.nat;
  c.dkernel.resize(nat);
  c.qkernel.resize(nat);
  c.mrad.resize(nat);
  c.dmrdcn.resize(nat);
  for (int i = 0; i < nat; i++) {
    int iz = m.z[i] - 1;
    c.dkernel[i] = GFN2_P_DKERNEL[iz];
    c.qkernel[i] = GFN2_P_QKERNEL[iz];
    double arg = cn[i] - GFN2_P_VCN[iz] - 1.2;
    double t1 = std::exp(-4.0 * arg);
    double t2 = (5.0 - GFN2_P_RAD[iz]) / (1.0 + t1);
    c.mrad[i] = GFN2_P_RAD[iz] + t2;
    c.dmrdcn[i] = -t2 * 4.0 * t1 / (1.0 + t1);
  }
  return c;
}

// g3, g5 of the AES pair (coulomb.rs:383-394)
inline void aes_g(const Model& m, const Coulomb& c, int i, int j, double& g3, double& g5) {
  double v[3];
  for (int k = 0; k < 3; k++) v[k] = m.xyz[3 * i + k] - m.xyz[3 * j + k];
  double r1 = std::sqrt(dot3(v, v));
  double g1 = 1.0 / r1;
  double g3v = g1 * g1 * g1, g5v = g3v * g1 * g1;
  double rr = 0.5 * (c.mrad[j] + c.mrad[i]) * g1;
  double f3 = 1.0 / (1.0 + 6.0 * std::pow(rr, 3.0));
  double f5 = 1.0 / (1.0 + 6.0 * std::pow(rr, 4.0));
  g3 = g3v * f3;
  g5 = g5v * f5;
}

struct Moments {
  vec qsh, qat, dp, qp;  // dp [nat][3], qp [nat][6]
};
struct Pots {
  vec vsh, vat, vdp, vqp;
};

// coulomb.rs:407-502 AES potentials (added to vat, vdp, vqp)
void aes_potentials(const Model& m, const Coulomb& c, const Moments& s, Pots& p) {
  int nat = m.nat;
  for (int a = 0; a < nat; a++) {
    for (int b = 0; b < nat; b++) {
      if (a == b) continue;
      double g3, g5;
      aes_g(m, c, std::max(a, b), std::min(a, b), g3, g5);
      double w[3];
      for (int k = 0; k < 3; k++) w[k] = m.xyz[3 * a + k] - m.xyz[3 * b + k];
      double r2 = dot3(w, w);
      double tc[6] = {w[0] * w[0], w[0] * w[1], w[1] * w[1], w[0] * w[2], w[1] * w[2], w[2] * w[2]};
      double t[6];
      for (int k = 0; k < 6; k++) t[k] = QSCALE[k] * tc[k] * g5;
      const double* mb = &s.dp[3 * b];
      double mbw = dot3(mb, w);
      p.vat[a] += mbw * g3;
      for (int k = 0; k < 3; k++) p.vdp[3 * a + k] -= w[k] * g3 * s.qat[b];
      for (int k = 0; k < 3; k++) p.vdp[3 * a + k] += g5 * (r2 * mb[k] - 3.0 * w[k] * mbw);
      double thb = 0.0;
      for (int k = 0; k < 6; k++) {
        p.vqp[6 * a + k] += t[k] * s.qat[b];
        thb += t[k] * s.qp[6 * b + k];
      }
      p.vat[a] += thb;
    }
  }
  for (int i = 0; i < nat; i++) {
    for (int k = 0; k < 3; k++) p.vdp[3 * i + k] += 2.0 * c.dkernel[i] * s.dp[3 * i + k];
    for (int k = 0; k < 6; k++) p.vqp[6 * i + k] += 2.0 * c.qkernel[i] * s.qp[6 * i + k] * QSCALE[k];
  }
}

// coulomb.rs:668-731 AES atom energies
void aes_energies(const Model& m, const Coulomb& c, const Moments& s, vec& e) {
  int nat = m.nat;
  for (int i = 1; i < nat; i++)
    for (int j = 0; j < i; j++) {
      double g3, g5;
      aes_g(m, c, i, j, g3, g5);
      double v[3];
      for (int k = 0; k < 3; k++) v[k] = m.xyz[3 * i + k] - m.xyz[3 * j + k];
      double r2 = dot3(v, v);
      const double *di = &s.dp[3 * i], *dj = &s.dp[3 * j];
      double miv = dot3(di, v), mjv = dot3(dj, v);
      double en = (s.qat[i] * mjv - s.qat[j] * miv) * g3;
      en += g5 * (r2 * dot3(di, dj) - 3.0 * miv * mjv);
      double tc[6] = {v[0] * v[0], v[0] * v[1], v[1] * v[1], v[0] * v[2], v[1] * v[2], v[2] * v[2]};
      for (int k = 0; k < 6; k++) en += QSCALE[k] * tc[k] * g5 * (s.qat[j] * s.qp[6 * i + k] + s.qat[i] * s.qp[6 * j + k]);
      e[i] += 0.5 * en;
      e[j] += 0.5 * en;
    }
  for (int i = 0; i < nat; i++) {
    e[i] += c.dkernel[i] * dot3(&s.dp[3 * i], &s.dp[3 * i]);
    double qq = 0.0;
    for (int k = 0; k < 6; k++) qq += QSCALE[k] * s.qp[6 * i + k] * s.qp[6 * i + k];
    e[i] += c.qkernel[i] * qq;
  }
}

// coulomb.rs:896-1001 AES gradient; adds dE/dCN to dedcn
void aes_gradient(const Model& m, const Coulomb& c, const Moments& s, vec& g, vec& dedcn) {
  int nat = m.nat;
  vec dedr(nat, 0.0);
  const double kdmp3 = 3.0, kdmp5 = 4.0;
  for (int iat = 1; iat < nat; iat++)
    for (int jat = 0; jat < iat; jat++) {
      double rr = 0.5 * (c.mrad[iat] + c.mrad[jat]);
      double vec[3];
      for (int k = 0; k < 3; k++) vec[k] = m.xyz[3 * jat + k] - m.xyz[3 * iat + k];
      double r2 = dot3(vec, vec), r1 = std::sqrt(r2);
      double g1 = 1.0 / r1, g3 = g1 * g1 * g1, g5 = g3 * g1 * g1, g7 = g5 * g1 * g1;
      double fdmp3 = 1.0 / (1.0 + 6.0 * std::pow(rr * g1, kdmp3));
      double ddmp3 = -3.0 * g5 * fdmp3 - kdmp3 * fdmp3 * (fdmp3 - 1.0) * g5;
      double fdmp5 = 1.0 / (1.0 + 6.0 * std::pow(rr * g1, kdmp5));
      double ddmp5 = -5.0 * fdmp5 - kdmp5 * (fdmp5 * fdmp5 - fdmp5);
      double qi = s.qat[iat], qj = s.qat[jat];
      const double *di = &s.dp[3 * iat], *dj = &s.dp[3 * jat];
      double dpiqj = dot3(vec, di) * qj, qidpj = dot3(vec, dj) * qi;
      double fddr = 3.0 * (dpiqj - qidpj) * kdmp3 * fdmp3 * g3 * (fdmp3 / rr) * std::pow(rr * g1, kdmp3);
      double dg[3];
      for (int k = 0; k < 3; k++) dg[k] = -ddmp3 * vec[k] * (dpiqj - qidpj) + fdmp3 * g3 * (qi * dj[k] - qj * di[k]);
      double dpidpj = dot3(dj, di), dpiv = dot3(di, vec), dpjv = dot3(dj, vec);
      double edd = dpidpj * r2 - 3.0 * dpjv * dpiv;
      fddr += 3.0 * edd * kdmp5 * fdmp5 * g5 * (fdmp5 / rr) * std::pow(rr * g1, kdmp5);
      for (int k = 0; k < 3; k++)
        dg[k] += -2.0 * fdmp5 * g5 * dpidpj * vec[k] + 3.0 * fdmp5 * g5 * (dpiv * dj[k] + dpjv * di[k]) -
                 edd * ddmp5 * g7 * vec[k];
      const double *ti = &s.qp[6 * iat], *tj = &s.qp[6 * jat];
      double eq = 2.0 * (qj * ti[1] + tj[1] * qi) * vec[0] * vec[1] + 2.0 * (qj * ti[3] + tj[3] * qi) * vec[0] * vec[2] +
                  2.0 * (qj * ti[4] + tj[4] * qi) * vec[1] * vec[2] + (qj * ti[0] + tj[0] * qi) * vec[0] * vec[0] +
                  (qj * ti[2] + tj[2] * qi) * vec[1] * vec[1] + (qj * ti[5] + tj[5] * qi) * vec[2] * vec[2];
      fddr += eq * 3.0 * kdmp5 * fdmp5 * g5 * fdmp5 / rr * std::pow(rr * g1, kdmp5);
      double tvj[3] = {vec[0] * tj[0] + vec[1] * tj[1] + vec[2] * tj[3], vec[0] * tj[1] + vec[1] * tj[2] + vec[2] * tj[4],
                       vec[0] * tj[3] + vec[1] * tj[4] + vec[2] * tj[5]};
      double tvi[3] = {vec[0] * ti[0] + vec[1] * ti[1] + vec[2] * ti[3], vec[0] * ti[1] + vec[1] * ti[2] + vec[2] * ti[4],
                       vec[0] * ti[3] + vec[1] * ti[4] + vec[2] * ti[5]};
      for (int k = 0; k < 3; k++)
        dg[k] += -eq * ddmp5 * g7 * vec[k] - 2.0 * fdmp5 * g5 * qi * tvj[k] - 2.0 * fdmp5 * g5 * qj * tvi[k];
      dedr[iat] += fddr;
      dedr[jat] += fddr;
      for (int k = 0; k < 3; k++) {
        g[3 * iat + k] += dg[k];
        g[3 * jat + k] -= dg[k];
      }
    }
  for (int i = 0; i < nat; i++) dedcn[i] += dedr[i] * c.dmrdcn[i];
}

// coulomb.rs:123-169 ES2 gradient
void es2_gradient(const Model& m, const Coulomb& c, const vec& qsh, vec& g) {
  for (int iat = 1; iat < m.nat; iat++)
    for (int jat = 0; jat < iat; jat++) {
      double v[3];
      for (int k = 0; k < 3; k++) v[k] = m.xyz[3 * iat + k] - m.xyz[3 * jat + k];
      double r1 = std::sqrt(dot3(v, v));
      double dg = 0.0;
      for (int ish = 0; ish < m.nsh_at[iat]; ish++)
        for (int jsh = 0; jsh < m.nsh_at[jat]; jsh++) {
          int a = m.ish_at[iat] + ish, b = m.ish_at[jat] + jsh;
          double eta = 0.5 * (c.hardness[a] + c.hardness[b]);
          double dtmp = 1.0 / (std::pow(r1, 2.0) + std::pow(eta, -2.0));
          dtmp = -std::pow(r1, 0.0) * dtmp * std::pow(dtmp, 0.5);
          dg += dtmp * qsh[a] * qsh[b];
        }
      for (int k = 0; k < 3; k++) {
        g[3 * iat + k] += dg * v[k];
        g[3 * jat + k] -= dg * v[k];
      }
    }
}

// ------------------------------------------------------------------ repulsion (repulsion.rs:44-73)
double repulsion(const Model& m, vec& e, vec* g) {
  double c2 = m.rep_cut2;
  for (int i = 0; i < m.nat; i++)
    for (int j = 0; j < i; j++) {
      double v[3];
      for (int k = 0; k < 3; k++) v[k] = m.xyz[3 * i + k] - m.xyz[3 * j + k];
      double r2 = dot3(v, v);
      if (r2 > c2 || r2 < 1e-12) continue;
      int zi = m.z[i] - 1, zj = m.z[j] - 1;
      double r1 = std::sqrt(r2);
      double kexp = (m.z[i] > 2 || m.z[j] > 2) ? 1.5 : 1.0;
      double alpha = std::sqrt(GFN2_REP_ALPHA[zj] * GFN2_REP_ALPHA[zi]);
      double zeff = GFN2_REP_ZEFF[zj] * GFN2_REP_ZEFF[zi];
      double r1k = std::pow(r1, kexp);
      double exa = std::exp(-alpha * r1k);
      double de = zeff * exa / std::pow(r1, 1.0);
      e[i] += 0.5 * de;
      e[j] += 0.5 * de;
      if (g) {
        double f = -(alpha * r1k * kexp + 1.0) * de / r2;
        for (int k = 0; k < 3; k++) {
          (*g)[3 * i + k] += f * v[k];
          (*g)[3 * j + k] -= f * v[k];
        }
      }
    }
  double t = 0;
  for (double x : e) t += x;
  return t;
}

// ------------------------------------------------------------------ D4 (crates/d4/src)
const int MAXREF = AXTB_D4_MAX_REF, NFREQ = AXTB_D4_NFREQ;
const double FREQ[NFREQ] = {0.000001, 0.050000, 0.100000, 0.200000, 0.300000, 0.400000, 0.500000, 0.600000,
                            0.700000, 0.800000, 0.900000, 1.000000, 1.200000, 1.400000, 1.600000, 1.800000,
                            2.000000, 2.500000, 3.000000, 4.000000, 5.000000, 7.500000, 10.00000};

struct D4Options {
  int charge_model = 1;   // 0 EEQ reference charges, 1 GFN2
  int rcov_mode = 1;      // 0 mctc 4/3 radii + D4 EN table, 1 axtb codata-2022 radii + Pauling EN
  double cn_cut = 40.0, disp2 = 72.0, disp3 = 40.0, width2 = 1.0, width3 = 1.0;
};

// D4 rational damping of GFN2 (d4/src/disp.rs:20-22)
const double S6 = 1.0, S8 = 2.7, S9 = 5.0, A1 = 0.52, A2 = 5.0, ALP = 16.0;
const double GA = 3.0, GC = 2.0, WF = 6.0;

struct D4 {
  int nat = 0;
  vector<int> z;
  vec xyz;
  D4Options o;
  vector<int> nref;
  vector<vector<int>> ngw;
  vector<vec> refcn, refq;           // per atom [MAXREF]
  vector<vector<vec>> aiw;           // per atom [ref][freq]
  vec rcov, en, zeff, eta, r4r2;     // per atom
  vec cn;
};

// d4/src/model.rs:83-90 zeta function
inline double zeta_f(double a, double c, double qref, double qmod) {
  if (qmod < 0.0) return std::exp(a);
  return std::exp(a * (1.0 - std::exp(c * (1.0 - qref / qmod))));
}
inline void zeta_df(double a, double c, double qref, double qmod, double& z, double& dz) {
  if (qmod < 0.0) {
    z = std::exp(a);
    dz = 0.0;
    return;
  }
  double e = std::exp(c * (1.0 - qref / qmod));
  z = std::exp(a * (1.0 - e));
  dz = -((a * c * e * z * qref) / (qmod * qmod));
}

double trapzd(const double* p) {  // d4/src/model.rs:120-141
  double w[NFREQ];
  w[0] = 0.5 * (FREQ[1] - FREQ[0]);
  for (int k = 1; k < NFREQ - 1; k++) w[k] = 0.5 * ((FREQ[k] - FREQ[k - 1]) + (FREQ[k + 1] - FREQ[k]));
  w[NFREQ - 1] = 0.5 * (FREQ[NFREQ - 1] - FREQ[NFREQ - 2]);
  double s = 0.0;
  for (int k = 0; k < NFREQ; k++) s += p[k] * w[k];
  return s;
}

inline double tabz(const double* t, int z) { return (z == 0 || z > 118) ? 0.0 : t[z - 1]; }

D4 build_d4(int nat, const int* z, const double* xyz, const D4Options& o) {
  D4 d;
  d.nat = nat;
  d.z.assign(z, z + nat);
  d.xyz.assign(xyz, xyz + 3 * nat);
  d.o = o;
  for (int i = 0; i < nat; i++) {
    int Z = z[i];
    // d4/src/model.rs:60-71: covalent radius (mctc bohr) and r4r2
    double rc = o.rcov_mode ? covalent_rad_d3_codata2022(Z) : 4.0 / 3.0 * (AXTB_MCTC_AATOAU * EL_COVALENT_RAD_2009_AA[Z - 1]);
    d.rcov.push_back(rc);
    d.en.push_back(EL_PAULING_EN[Z - 1]);
    d.zeff.push_back(tabz(D4_EFFECTIVE_NUCLEAR_CHARGE, Z));
    d.eta.push_back(tabz(D4_CHEMICAL_HARDNESS, Z));
    d.r4r2.push_back(std::sqrt(0.5 * (D4_R4_OVER_R2[Z - 1] * std::sqrt((double)Z))));
    int nr = D4_REFN[Z - 1];
    d.nref.push_back(nr);
    // reference_ngw (model.rs:143-155)
    int cnc[20] = {0};
    cnc[0] = 1;
    auto icn = [](double x) {
      long v = std::lround(x);
      if (v < 0) v = 0;
      if (v > 19) v = 19;
      return (int)v;
    };
    for (int r = 0; r < nr; r++) cnc[icn(D4_REFCN[Z - 1][r])]++;
    vector<int> ngw(MAXREF, 1);
    for (int r = 0; r < nr; r++) {
      int c = cnc[icn(D4_REFCN[Z - 1][r])];
      ngw[r] = c * (c + 1) / 2;
    }
    d.ngw.push_back(ngw);
    vec rcn(MAXREF, 0.0), rq(MAXREF, 0.0);
    for (int r = 0; r < nr; r++) {
      rcn[r] = D4_REFCOVCN[Z - 1][r];
      rq[r] = o.charge_model == 0 ? D4_REFQ_EEQ[Z - 1][r] : D4_REFQ_GFN2[Z - 1][r];
    }
    d.refcn.push_back(rcn);
    d.refq.push_back(rq);
    // reference_alpha (model.rs:197-228)
    vector<vec> al(MAXREF, vec(NFREQ, 0.0));
    for (int r = 0; r < nr; r++) {
      int is = D4_REFSYS[Z - 1][r];
      if (is == 0) continue;
      double iz = tabz(D4_EFFECTIVE_NUCLEAR_CHARGE, is);
      double gam = tabz(D4_CHEMICAL_HARDNESS, is) * GC;
      double refh = o.charge_model == 0 ? D4_REFH_EEQ[Z - 1][r] : D4_REFH_GFN2[Z - 1][r];
      double zv = zeta_f(GA, gam, iz, refh + iz);
      double ascale = D4_ASCALE[Z - 1][r], hcount = D4_HCOUNT[Z - 1][r];
      for (int k = 0; k < NFREQ; k++) {
        double aiw = D4_SSCALE[is - 1] * D4_SECAIW[is - 1][k] * zv;
        double v = ascale * (D4_ALPHAIW[Z - 1][r][k] - hcount * aiw);
        al[r][k] = v > 0.0 ? v : 0.0;
      }
    }
    d.aiw.push_back(al);
  }
  if (o.rcov_mode == 0) {
    // dftd4 default: en table of dftd4 = Pauling (same values in mctc-lib)
  }
  // D4 CN (d4/src/ncoord.rs): erf counting with the EN factor, cutoff, Jacobian gradient order
  d.cn.assign(nat, 0.0);
  double c2 = o.cn_cut * o.cn_cut;
  for (int i = 0; i < nat; i++)
    for (int j = 0; j < i; j++) {
      double v[3];
      for (int k = 0; k < 3; k++) v[k] = xyz[3 * i + k] - xyz[3 * j + k];
      double r2 = dot3(v, v);
      if (r2 > c2 || r2 < 1e-12) continue;
      double x = std::fabs(d.en[i] - d.en[j]) + 19.08857;
      double w = 4.10451 * std::exp(-((x * x) / (2.0 * (11.28174 * 11.28174))));
      double c = w * erf_count(7.5, std::sqrt(r2), d.rcov[i] + d.rcov[j]);
      d.cn[i] += c;
      d.cn[j] += c;
    }
  return d;
}

inline double d4_refc6(const D4& d, int i, int j, int ri, int rj) {
  double prod[NFREQ];
  for (int k = 0; k < NFREQ; k++) prod[k] = d.aiw[i][ri][k] * d.aiw[j][rj][k];
  return 3.0 / PI * trapzd(prod);
}

// d4/src/model.rs:468-545 reference weights (and derivatives)
void d4_weights(const D4& d, const vec& q, vector<vec>& gw, vector<vec>* dgwdcn, vector<vec>* dgwdq) {
  int nat = d.nat;
  gw.assign(nat, vec(MAXREF, 0.0));
  if (dgwdcn) dgwdcn->assign(nat, vec(MAXREF, 0.0));
  if (dgwdq) dgwdq->assign(nat, vec(MAXREF, 0.0));
  double eps_norm = std::sqrt(2.2250738585072014e-308);
  for (int i = 0; i < nat; i++) {
    int nref = d.nref[i];
    double zi = d.zeff[i], gi = d.eta[i] * GC, cni = d.cn[i];
    double norm = 0.0, dnorm = 0.0;
    for (int r = 0; r < nref; r++)
      for (int g = 1; g <= d.ngw[i][r]; g++) {
        double wf = g * WF;
        double dd = cni - d.refcn[i][r];
        double gg = std::exp(-(wf * (dd * dd)));
        norm += gg;
        dnorm += 2.0 * wf * (d.refcn[i][r] - cni) * gg;
      }
    norm = std::fabs(norm) > eps_norm ? 1.0 / norm : 0.0;
    double maxcn = -INFINITY;
    for (int r = 0; r < nref; r++) maxcn = std::max(maxcn, d.refcn[i][r]);
    for (int r = 0; r < nref; r++) {
      double expw = 0.0, expd = 0.0;
      for (int g = 1; g <= d.ngw[i][r]; g++) {
        double wf = g * WF;
        double dd = cni - d.refcn[i][r];
        double gg = std::exp(-(wf * (dd * dd)));
        expw += gg;
        expd += 2.0 * wf * (d.refcn[i][r] - cni) * gg;
      }
      double gwk = expw * norm;
      if (!std::isfinite(gwk) || norm == 0.0) gwk = std::fabs(maxcn - d.refcn[i][r]) < 1e-12 ? 1.0 : 0.0;
      double qref = d.refq[i][r] + zi, qmod = q[i] + zi, zz, dz;
      zeta_df(GA, gi, qref, qmod, zz, dz);
      gw[i][r] = gwk * zz;
      if (dgwdq) (*dgwdq)[i][r] = gwk * dz;
      if (dgwdcn) {
        double dgwk = norm * (expd - expw * dnorm * norm);
        if (!std::isfinite(dgwk) || norm == 0.0) dgwk = 0.0;
        (*dgwdcn)[i][r] = dgwk * zz;
      }
    }
  }
}

// c6 and its derivatives from two weight sets (model.rs:607-627)
void d4_pair_c6(const D4& d, int i, int j, const vec& wi, const vec& di, const vec& wj, const vec& dj, double& c6,
                double& dci, double& dcj) {
  c6 = dci = dcj = 0.0;
  for (int ri = 0; ri < d.nref[i]; ri++)
    for (int rj = 0; rj < d.nref[j]; rj++) {
      double r = d4_refc6(d, i, j, ri, rj);
      c6 += wi[ri] * wj[rj] * r;
      dci += di[ri] * wj[rj] * r;
      dcj += wi[ri] * dj[rj] * r;
    }
}

// d4/src/disp.rs:43-63 smooth cutoff
inline void smooth_cutoff(double r, double cutoff, double width, double& sw, double& dsw) {
  if (width <= 0.0 || cutoff <= 0.0) {
    sw = 1.0;
    dsw = 0.0;
    return;
  }
  double w = std::min(width, cutoff), inner = cutoff - w;
  if (r <= inner) {
    sw = 1.0;
    dsw = 0.0;
  } else if (r >= cutoff) {
    sw = 0.0;
    dsw = 0.0;
  } else {
    double x = (cutoff - r) / w, x2 = x * x;
    sw = (x2 * x) * (10.0 + x * (-15.0 + 6.0 * x));
    double omx = 1.0 - x;
    dsw = -((30.0 * x2 * (omx * omx)) / w);
  }
}

inline double pw3(double x) { return (x * x) * x; }
inline double pw4(double x) { double x2 = x * x; return x2 * x2; }
inline double pw6(double x) { double x3 = pw3(x); return x3 * x3; }
inline double pw8(double x) { double x4 = pw4(x); return x4 * x4; }

// pair_e, pair_g of one pair (disp.rs:240-272)
inline bool d4_pair(const D4& d, int i, int j, double& e, double& gdr) {
  double v[3];
  for (int k = 0; k < 3; k++) v[k] = d.xyz[3 * i + k] - d.xyz[3 * j + k];
  double r2 = dot3(v, v);
  double cut = d.o.disp2;
  if (r2 > cut * cut || r2 < 2.220446049250313e-16) return false;
  double r = std::sqrt(r2), sw, dsw;
  smooth_cutoff(r, cut, d.o.width2, sw, dsw);
  if (sw <= 0.0) return false;
  double rr = 3.0 * d.r4r2[i] * d.r4r2[j];
  double r0 = A1 * std::sqrt(rr) + A2;
  double t6 = 1.0 / (pw3(r2) + pw6(r0)), t8 = 1.0 / (pw4(r2) + pw8(r0));
  double d6 = -(6.0 * (r2 * r2) * (t6 * t6)), d8 = -(8.0 * pw3(r2) * (t8 * t8));
  double e0 = S6 * t6 + S8 * rr * t8, g0 = S6 * d6 + S8 * rr * d8;
  e = sw * e0;
  gdr = sw * g0 + dsw * e0 / r;
  return true;
}

// two-body energy, potential dE/dq (disp.rs:380-455)
void d4_energy_potential(const D4& d, const vec& q, vec* e, vec* vat) {
  vector<vec> gw, dgwdq;
  d4_weights(d, q, gw, nullptr, &dgwdq);
  for (int i = 0; i < d.nat; i++)
    for (int j = 0; j < i; j++) {
      double pe, pg;
      if (!d4_pair(d, i, j, pe, pg)) continue;
      double c6, dqi, dqj;
      d4_pair_c6(d, i, j, gw[i], dgwdq[i], gw[j], dgwdq[j], c6, dqi, dqj);
      if (vat) {
        (*vat)[i] -= dqi * pe;
        (*vat)[j] -= dqj * pe;
      }
      if (e) {
        double de = -(c6 * pe * 0.5);
        (*e)[i] += de;
        (*e)[j] += de;
      }
    }
}

void d4_cn_gradient(const D4& d, const vec& dedcn, vec& g) {  // d4/src/ncoord.rs (Jacobian order)
  double c2 = d.o.cn_cut * d.o.cn_cut;
  for (int i = 0; i < d.nat; i++)
    for (int j = 0; j < i; j++) {
      double s = dedcn[i] + dedcn[j];
      if (s == 0.0) continue;
      double v[3];
      for (int k = 0; k < 3; k++) v[k] = d.xyz[3 * i + k] - d.xyz[3 * j + k];
      double r2 = dot3(v, v);
      if (r2 > c2 || r2 < 1e-12) continue;
      double r = std::sqrt(r2);
      double x = std::fabs(d.en[i] - d.en[j]) + 19.08857;
      double w = 4.10451 * std::exp(-((x * x) / (2.0 * (11.28174 * 11.28174))));
      double dc = w * erf_dcount(7.5, r, d.rcov[i] + d.rcov[j]);
      for (int k = 0; k < 3; k++) {
        double t = dc * v[k] / r * s;
        g[3 * i + k] += t;
        g[3 * j + k] -= t;
      }
    }
}

// disp.rs:460-484 two-body gradient at fixed charges
void d4_gradient(const D4& d, const vec& q, vec& g) {
  vector<vec> gw, dgwdcn;
  d4_weights(d, q, gw, &dgwdcn, nullptr);
  vec dedcn(d.nat, 0.0);
  for (int i = 0; i < d.nat; i++)
    for (int j = 0; j < i; j++) {
      double pe, pg;
      if (!d4_pair(d, i, j, pe, pg)) continue;
      double c6, dci, dcj;
      d4_pair_c6(d, i, j, gw[i], dgwdcn[i], gw[j], dgwdcn[j], c6, dci, dcj);
      dedcn[i] -= dci * pe;
      dedcn[j] -= dcj * pe;
      for (int k = 0; k < 3; k++) {
        double dg = -c6 * pg * (d.xyz[3 * i + k] - d.xyz[3 * j + k]);
        g[3 * i + k] += dg;
        g[3 * j + k] -= dg;
      }
    }
  d4_cn_gradient(d, dedcn, g);
}

// ATM three-body at q = 0, per triple (disp.rs:665-870; the pair-sum form gives the same
// values up to rounding). Adds atom energies and, with g, the gradient incl. the D4 CN chain.
void d4_three_body(const D4& d, vec& e_out, vec* g) {
  int nat = d.nat;
  vec q0(nat, 0.0);
  vector<vec> gw, dgwdcn;
  d4_weights(d, q0, gw, &dgwdcn, nullptr);
  double cut = d.o.disp3, alp3 = ALP / 3.0;
  // pair table
  Mat C6(nat, nat), QA(nat, nat), SW(nat, nat), DSWR(nat, nat), T(nat, nat), R2(nat, nat);
  vector<char> ok((size_t)nat * nat, 0);
  for (int a = 0; a < nat; a++)
    for (int b = 0; b < a; b++) {
      double v[3];
      for (int k = 0; k < 3; k++) v[k] = d.xyz[3 * b + k] - d.xyz[3 * a + k];
      double r2 = dot3(v, v);
      if (r2 > cut * cut || r2 < 2.220446049250313e-16) continue;
      double r = std::sqrt(r2), sw, dsw;
      smooth_cutoff(r, cut, d.o.width3, sw, dsw);
      double rr = 3.0 * d.r4r2[a] * d.r4r2[b];
      double r0 = A1 * std::sqrt(rr) + A2;
      double c6, dca, dcb;
      d4_pair_c6(d, a, b, gw[a], dgwdcn[a], gw[b], dgwdcn[b], c6, dca, dcb);
      double t = std::pow(r0 / r, alp3);
      double qa = c6 != 0.0 ? dca / c6 : 0.0, qb = c6 != 0.0 ? dcb / c6 : 0.0;
      for (int pass = 0; pass < 2; pass++) {
        int x = pass ? b : a, y = pass ? a : b;
        ok[(size_t)y * nat + x] = sw > 0.0;
        C6(x, y) = c6;
        SW(x, y) = sw;
        DSWR(x, y) = dsw / r;
        T(x, y) = t;
        R2(x, y) = r2;
        QA(x, y) = pass ? qb : qa;  // dlnC6/dCN of the first index
      }
    }
  vec dedcn(nat, 0.0);
  // triples (i > j > k) by i over the threads, each with its own sums, added in thread order
  const int nt = std::max(1, omp_get_max_threads());
  vector<vec> te(nt, vec(nat, 0.0)), tg(nt, vec(g ? 3 * nat : 0, 0.0)), tc(nt, vec(nat, 0.0));
#pragma omp parallel
  {
  const int tid = omp_get_thread_num();
  vec& e = te[tid];
  vec* gt = g ? &tg[tid] : nullptr;
  vec& dedcn = tc[tid];
#pragma omp for schedule(static, 1)
  for (int i = 0; i < nat; i++)
    for (int j = 0; j < i; j++) {
      if (!ok[(size_t)j * nat + i]) continue;
      for (int k = 0; k < j; k++) {
        if (!ok[(size_t)k * nat + i] || !ok[(size_t)k * nat + j]) continue;
        double swij = SW(i, j), swik = SW(i, k), swjk = SW(j, k);
        double sw = swij * swik * swjk;
        if (sw <= 0.0) continue;
        double r2ij = R2(i, j), r2ik = R2(i, k), r2jk = R2(j, k);
        double c9 = -(S9 * std::sqrt(std::fabs(C6(i, j) * C6(i, k) * C6(j, k))));
        double ir2 = 1.0 / r2ij * (1.0 / r2ik) * (1.0 / r2jk);
        double ir1 = std::sqrt(ir2);
        double ir3 = ir2 * ir1, ir5 = ir3 * ir2;
        double rpow = T(i, j) * T(i, k) * T(j, k);
        double fdmp = 1.0 / (1.0 + 6.0 * rpow);
        double ang = 0.375 * (r2ij + r2jk - r2ik) * (r2ij - r2jk + r2ik) * (-r2ij + r2jk + r2ik) * ir5 + ir3;
        double rr = ang * fdmp;
        double de0 = rr * c9;
        double de = de0 * sw;
        e[i] -= de / 3.0;
        e[j] -= de / 3.0;
        e[k] -= de / 3.0;
        if (!gt) continue;
        double dfdmp = -(2.0 * ALP * rpow * (fdmp * fdmp));
        double swc9 = sw * c9;
        auto dangf = [&](double a, double b, double c) {  // derivative of ang wrt a = r2 of the pair
          return -((0.375 * (pw3(a) + (a * a) * (b + c) + a * (3.0 * (b * b) + 2.0 * b * c + 3.0 * (c * c)) -
                             5.0 * ((b - c) * (b - c)) * (b + c))) * ir5);
        };
        double fij = swc9 * (-(dangf(r2ij, r2jk, r2ik) * fdmp) + ang * dfdmp) / r2ij - de0 * DSWR(i, j) * swik * swjk;
        double fik = swc9 * (-(dangf(r2ik, r2jk, r2ij) * fdmp) + ang * dfdmp) / r2ik - de0 * DSWR(i, k) * swij * swjk;
        double fjk = swc9 * (-(dangf(r2jk, r2ik, r2ij) * fdmp) + ang * dfdmp) / r2jk - de0 * DSWR(j, k) * swij * swik;
        for (int x = 0; x < 3; x++) {
          double vij = d.xyz[3 * j + x] - d.xyz[3 * i + x];
          double vik = d.xyz[3 * k + x] - d.xyz[3 * i + x];
          double vjk = d.xyz[3 * k + x] - d.xyz[3 * j + x];
          double dgij = fij * vij, dgik = fik * vik, dgjk = fjk * vjk;
          (*gt)[3 * i + x] -= dgij + dgik;
          (*gt)[3 * j + x] += dgij - dgjk;
          (*gt)[3 * k + x] += dgik + dgjk;
        }
        double h = de * 0.5;
        dedcn[i] -= h * (QA(i, j) + QA(i, k));
        dedcn[j] -= h * (QA(j, i) + QA(j, k));
        dedcn[k] -= h * (QA(k, i) + QA(k, j));
      }
    }
  }
  for (int t = 0; t < nt; t++) {
    for (int a = 0; a < nat; a++) {
      e_out[a] += te[t][a];
      dedcn[a] += tc[t][a];
    }
    if (g)
      for (int a = 0; a < 3 * nat; a++) (*g)[a] += tg[t][a];
  }
  if (g) d4_cn_gradient(d, dedcn, *g);
}

// ------------------------------------------------------------------ eigensolver
// Householder reduction to tridiagonal form with accumulated transformation, then implicit QL
// with Wilkinson-type shifts (the textbook EISPACK tred2/tql2 scheme, own implementation).
// In: a (n x n, column-major, symmetric; overwritten by the eigenvectors). Out: w ascending.
void eigh(int n, vec& a, vec& w) {
  // the matrix is symmetric, so its column-major storage read row by row is the same matrix:
  // R(i, j) = a[i n + j] keeps every inner loop below on contiguous memory
  auto R = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
  vec d(n), e(n), u(n);
  const int nt = std::max(1, omp_get_max_threads());
  vector<vec> part(nt, vec(n));
  // Householder reduction, last row first: row i (k <= l) becomes the reflector u, the leading
  // (l + 1) x (l + 1) lower triangle gets A - u p^T - p u^T
  for (int i = n - 1; i > 0; i--) {
    const int l = i - 1;
    double h = 0.0, scale = 0.0;
    double* ri = &a[(size_t)i * n];
    if (l > 0) {
      for (int k = 0; k <= l; k++) scale += std::fabs(ri[k]);
      if (scale == 0.0) {
        e[i] = ri[l];
      } else {
        for (int k = 0; k <= l; k++) {
          ri[k] /= scale;
          h += ri[k] * ri[k];
        }
        double f = ri[l];
        double g = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
        e[i] = scale * g;
        h -= f * g;
        ri[l] = f - g;
        // p = A u / h from the lower triangle: row j's dot with u, and row j's entries scattered
        // into the transposed products (per-thread sums, added in thread order)
#pragma omp parallel
        {
          const int t = omp_get_thread_num();
          double* acc = part[t].data();
          std::fill(acc, acc + l + 1, 0.0);
#pragma omp for schedule(static, 16)
          for (int j = 0; j <= l; j++) {
            const double* rj = &a[(size_t)j * n];
            const double uj = ri[j];
            double s = rj[j] * uj;
            for (int k = 0; k < j; k++) {
              s += rj[k] * ri[k];
              acc[k] += rj[k] * uj;
            }
            acc[j] += s;
          }
        }
        f = 0.0;
        for (int j = 0; j <= l; j++) {
          double s = 0.0;
          for (int t = 0; t < nt; t++) s += part[t][j];
          e[j] = s / h;
          f += e[j] * ri[j];
        }
        const double hh = f / (h + h);
        for (int j = 0; j <= l; j++) e[j] -= hh * ri[j];
#pragma omp parallel for schedule(static, 16)
        for (int j = 0; j <= l; j++) {
          double* rj = &a[(size_t)j * n];
          const double fj = ri[j], gj = e[j];
          for (int k = 0; k <= j; k++) rj[k] -= fj * e[k] + gj * ri[k];
        }
        for (int j = 0; j <= l; j++) R(j, i) = ri[j] / h;
      }
    } else {
      e[i] = ri[l];
    }
    d[i] = h;
  }
  d[0] = 0.0;
  e[0] = 0.0;
  // the accumulated transformation: Q = H_1 ... H_{n-1}, built up from the first reflector
  for (int i = 0; i < n; i++) {
    const int l = i - 1;
    if (d[i] != 0.0) {
      const double* ri = &a[(size_t)i * n];
      // g_j = sum_k R(i, k) R(k, j), then R(k, j) -= g_j R(k, i), for j, k <= l
#pragma omp parallel for schedule(static)
      for (int j0 = 0; j0 <= l; j0 += 64) {
        const int j1 = std::min(l + 1, j0 + 64);
        for (int j = j0; j < j1; j++) u[j] = 0.0;
        for (int k = 0; k <= l; k++) {
          const double rik = ri[k];
          const double* rk = &a[(size_t)k * n];
          for (int j = j0; j < j1; j++) u[j] += rik * rk[j];
        }
      }
#pragma omp parallel for schedule(static)
      for (int k = 0; k <= l; k++) {
        double* rk = &a[(size_t)k * n];
        const double rki = rk[i];
        for (int j = 0; j <= l; j++) rk[j] -= u[j] * rki;
      }
    }
    d[i] = R(i, i);
    R(i, i) = 1.0;
    for (int j = 0; j <= l; j++) R(j, i) = R(i, j) = 0.0;
  }
  // implicit QL on (d, e); the rotations of the vectors are recorded and applied in batches, row
  // by row (a row of R holds one component of every vector; the recurrence on d and e does not
  // read the vectors, so deferring them changes nothing)
  for (int i = 1; i < n; i++) e[i - 1] = e[i];
  if (n > 0) e[n - 1] = 0.0;
  vector<int> rot_i;
  vec rot_c, rot_s;
  auto flush = [&]() {
    const size_t nr = rot_i.size();
    if (!nr) return;
#pragma omp parallel for schedule(static)
    for (int k0 = 0; k0 < n; k0 += 8) {
      const int k1 = std::min(n, k0 + 8);
      for (size_t r = 0; r < nr; r++) {
        const int i = rot_i[r];
        const double c = rot_c[r], s = rot_s[r];
        for (int k = k0; k < k1; k++) {
          double* rk = &a[(size_t)k * n];
          const double f = rk[i + 1];
          rk[i + 1] = s * rk[i] + c * f;
          rk[i] = c * rk[i] - s * f;
        }
      }
    }
    rot_i.clear();
    rot_c.clear();
    rot_s.clear();
  };
  for (int l = 0; l < n; l++) {
    int iter = 0, mm;
    do {
      for (mm = l; mm < n - 1; mm++) {
        double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
        if (std::fabs(e[mm]) <= 2.220446049250313e-16 * dd) break;
      }
      if (mm != l) {
        if (iter++ == 60) throw Error(2, "QL did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[mm] - d[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = mm - 1; i >= l; i--) {
          double f = s * e[i], b = c * e[i];
          e[i + 1] = (r = std::hypot(f, g));
          if (r == 0.0) {
            d[i + 1] -= p;
            e[mm] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          d[i + 1] = g + (p = s * r);
          g = c * r - b;
          rot_i.push_back(i);
          rot_c.push_back(c);
          rot_s.push_back(s);
        }
        if (rot_i.size() >= ((size_t)1 << 17)) flush();
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[mm] = 0.0;
      }
    } while (mm != l);
  }
  flush();
  // sort ascending; eigenvector j is column idx[j] of R
  vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  vec out((size_t)n * n);
  w.resize(n);
  for (int j = 0; j < n; j++) w[j] = d[idx[j]];
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (int k = 0; k < n; k++) out[(size_t)j * n + k] = a[(size_t)k * n + idx[j]];
  a.swap(out);
}

// ------------------------------------------------------------------ occupations (base/src/occupations.rs)
int aufbau(double nel, vec& occ) {
  int n = (int)occ.size();
  int full = (int)std::floor(nel);
  for (int i = 0; i < n; i++) occ[i] = i < full ? 1.0 : 0.0;
  double rem = std::fmod(nel, 1.0);
  if (full < n) occ[full] = rem;
  return full + (rem > 0.5 ? 1 : 0);
}

inline void fermi(double arg, double kt, double& f, double& df) {
  if (arg > 50.0) {
    f = 0.0;
    df = 0.0;
  } else if (arg < -50.0) {
    f = 1.0;
    df = 0.0;
  } else {
    f = 1.0 / (std::exp(arg) + 1.0);
    df = -f * (1.0 - f) / kt;
  }
}

// occupations.rs:94-125 (Newton::XTB680: tol 1e5 eps, 200 steps)
double fill_tblite(double nel, double kt, const vec& emo, vec& occ) {
  int homo = aufbau(nel, occ);
  if (!(nel > 0.0 && kt > 0.0)) return NAN;
  int nao = (int)emo.size();
  double sqrttiny = std::sqrt(2.2250738585072014e-308);
  int i0 = std::max(homo, 1) - 1, i1 = std::min(homo + 1, nao) - 1;
  double ef = 0.5 * (emo[i0] + emo[i1]), ef_used = ef;
  for (int it = 0; it < 200; it++) {
    double total = 0.0, dtotal = 0.0;
    ef_used = ef;
    for (int i = 0; i < nao; i++) {
      double f, df;
      fermi((emo[i] - ef) / kt, kt, f, df);
      occ[i] = f;
      total += f;
      dtotal -= df;
    }
    if (std::fabs(nel - total) <= 1e5 * 2.220446049250313e-16) break;
    if (dtotal > sqrttiny)
      ef += (nel - total) / dtotal;
    else
      break;
  }
  return ef_used;
}

double entropy_tblite(const vec& occ, double kt) {  // occupations.rs:179-181
  double s = 0.0;
  for (double o : occ) s += std::log(std::pow(o, o) * std::pow(1.0 - o, 1.0 - o));
  return s * kt;
}

// ------------------------------------------------------------------ Broyden (scf/mixing.rs:192-338)
struct Broyden {
  int ndim, memory, iter = 0;
  double damp;
  vector<vec> df, u;
  vector<vec> a;
  vec omega, q_in, dq, qlast, dqlast;
  Broyden(int n, int mem, double d) : ndim(n), memory(std::max(mem, 1)), damp(d), q_in(n), dq(n), qlast(n), dqlast(n) {}
  void set_residual(const vec& out) {
    for (int k = 0; k < ndim; k++) dq[k] = out[k] - q_in[k];
  }
  static double dotv(const vec& x, const vec& y) {
    double s = 0;
    for (size_t i = 0; i < x.size(); i++) s += x[i] * y[i];
    return s;
  }
  bool next() {
    iter++;
    int n = ndim;
    double alpha = damp;
    if (iter == 1) {
      dqlast = dq;
      qlast = q_in;
      for (int i = 0; i < n; i++) q_in[i] += alpha * dq[i];
      return true;
    }
    int itn = iter - 1;
    int it1 = (itn - 1) % memory;
    int nhist = std::min(memory, itn);
    int first = itn + 1 > memory ? itn + 1 - memory : 1;
    vector<int> hist(nhist);
    for (int ih = 0; ih < nhist; ih++) hist[ih] = (first + ih - 1) % memory;
    const double OMEGA0 = 0.01, MINW = 1.0, MAXW = 100000.0, WFAC = 0.01;
    if (it1 == (int)df.size()) {
      df.push_back(vec(n));
      u.push_back(vec(n));
      omega.push_back(0.0);
      for (auto& row : a) row.push_back(0.0);
      a.push_back(vec(df.size(), 0.0));
    }
    double norm = std::sqrt(dotv(dq, dq));
    omega[it1] = norm > WFAC / MAXW ? WFAC / norm : MAXW;
    if (omega[it1] < MINW) omega[it1] = MINW;
    for (int i = 0; i < n; i++) df[it1][i] = dq[i] - dqlast[i];
    double inv = 1.0 / std::max(std::sqrt(dotv(df[it1], df[it1])), 2.220446049250313e-16);
    for (double& x : df[it1]) x *= inv;
    vec c(nhist);
    vector<vec> beta(nhist, vec(nhist));
    for (int ih = 0; ih < nhist; ih++) {
      int i = hist[ih];
      c[ih] = omega[i] * dotv(df[i], dq);
      for (int jh = 0; jh < nhist; jh++) {
        int j = hist[jh];
        a[i][j] = dotv(df[i], df[j]);
      }
    }
    for (int ih = 0; ih < nhist; ih++) {
      for (int jh = 0; jh < nhist; jh++) beta[ih][jh] = omega[hist[ih]] * omega[hist[jh]] * a[hist[ih]][hist[jh]];
      beta[ih][ih] += OMEGA0 * OMEGA0;
    }
    if (!gauss_last_max(beta, c)) return false;
    for (int k = 0; k < n; k++) u[it1][k] = alpha * df[it1][k] + inv * (q_in[k] - qlast[k]);
    dqlast = dq;
    qlast = q_in;
    for (int k = 0; k < n; k++) q_in[k] += alpha * dq[k];
    for (int ih = 0; ih < nhist; ih++) {
      int i = hist[ih];
      double f = omega[i] * c[ih];
      for (int k = 0; k < n; k++) q_in[k] -= f * u[i][k];
    }
    return true;
  }
};

// ------------------------------------------------------------------ DIIS (scf/mixing.rs:340-456)
// Pulay mixing with restarts: the last DIIS_SIZE inputs and residuals, coefficients by modified
// Gram-Schmidt on the differences to the newest residual (newest first, a nearly dependent one ends
// the basis), restart from the best iterate with half the damping on a stall or a divergence.
const int DIIS_SIZE = 30;                 // mixing.rs:341
const double DIIS_DEPENDENCE = 1e-2;      // mixing.rs:346
const int DIIS_STALL = 12;                // mixing.rs:349
const double DIIS_RESTART_GROWTH = 1e3;   // mixing.rs:352
const double DIIS_RESTART_FLOOR = 0.1;    // mixing.rs:354

inline double dot_seq(const vec& x, const vec& y) {  // base/src/math.rs:69-71
  double s = 0.0;
  for (size_t i = 0; i < x.size(); i++) s += x[i] * y[i];
  return s;
}

// mixing.rs:467-513
vec diis_coefficients(const std::deque<vec>& r) {
  const int m = (int)r.size();
  const vec& last = r[m - 1];
  const size_t n = last.size();
  vector<vec> basis, rcol;
  vector<int> kept;
  for (int j = m - 2; j >= 0; j--) {
    vec d(n);
    for (size_t k = 0; k < n; k++) d[k] = last[k] - r[j][k];
    const double norm0 = std::sqrt(dot_seq(d, d));
    if (std::isnan(norm0) || norm0 <= 0.0) break;
    vec col(basis.size() + 1, 0.0);
    for (size_t l = 0; l < basis.size(); l++) {
      const double h = dot_seq(basis[l], d);
      col[l] = h;
      for (size_t k = 0; k < n; k++) d[k] -= h * basis[l][k];
    }
    const double nr = std::sqrt(dot_seq(d, d));
    if (nr < DIIS_DEPENDENCE * norm0) break;
    col[basis.size()] = nr;
    for (double& x : d) x /= nr;
    basis.push_back(d);
    rcol.push_back(col);
    kept.push_back(j);
  }
  const int p = (int)basis.size();
  vec rhs(p), gl(p, 0.0);
  for (int l = 0; l < p; l++) rhs[l] = dot_seq(basis[l], last);
  for (int l = p - 1; l >= 0; l--) {
    double v = rhs[l];
    for (int k = l + 1; k < p; k++) v -= rcol[k][l] * gl[k];
    gl[l] = v / rcol[l][l];
  }
  vec c(m, 0.0);
  c[m - 1] = 1.0;
  for (int l = 0; l < p; l++) {
    c[kept[l]] += gl[l];
    c[m - 1] -= gl[l];
  }
  return c;
}

struct Diis {
  double damp;
  std::deque<vec> hq, hr;
  bool has_best = false;
  double best_norm = 0.0;
  vec best_q, best_r, q_in, dq;
  int since_best = 0;
  Diis(int n, double d) : damp(d), q_in(n), dq(n) {}
  // mixing.rs:410-448
  bool next() {
    const double norm = std::sqrt(dot_seq(dq, dq));
    if (!std::isfinite(norm)) return false;
    if (has_best) {
      const bool diverging = norm > DIIS_RESTART_GROWTH * best_norm && norm > DIIS_RESTART_FLOOR;
      if (diverging || since_best >= DIIS_STALL) {
        since_best = 0;
        hq.clear();
        hr.clear();
        damp *= 0.5;
        for (size_t k = 0; k < q_in.size(); k++) q_in[k] = best_q[k] + damp * best_r[k];
        return true;
      }
    }
    if (!has_best || norm < best_norm) {
      has_best = true;
      best_norm = norm;
      best_q = q_in;
      best_r = dq;
      since_best = 0;
    } else {
      since_best++;
    }
    if ((int)hq.size() == DIIS_SIZE) {
      hq.pop_front();
      hr.pop_front();
    }
    hq.push_back(q_in);
    hr.push_back(dq);
    const vec c = diis_coefficients(hr);
    for (size_t k = 0; k < q_in.size(); k++) {
      double x = 0.0;
      for (size_t i = 0; i < c.size(); i++)
        if (c[i] != 0.0) x += c[i] * (hq[i][k] + damp * hr[i][k]);
      q_in[k] = x;
    }
    return true;
  }
};

// The vector mixer of one SCF attempt: Broyden or DIIS (MixerKind, scf/mixing.rs:83-118).
enum MixerKind { MIX_BROYDEN = 0, MIX_DIIS = 1, MIX_AUTO = 2 };
struct VecMixer {
  int kind;
  Broyden br;
  Diis di;
  VecMixer(int k, int n, int memory, double damp) : kind(k), br(n, memory, damp), di(n, damp) {}
  vec& q_in() { return kind == MIX_DIIS ? di.q_in : br.q_in; }
  vec& dq() { return kind == MIX_DIIS ? di.dq : br.dq; }
  void set_residual(const vec& out) {
    vec& q = q_in();
    vec& d = dq();
    for (size_t k = 0; k < d.size(); k++) d[k] = out[k] - q[k];
  }
  bool next() { return kind == MIX_DIIS ? di.next() : br.next(); }
};

// ------------------------------------------------------------------ implicit solvation (crates/solvation)
// The model's setup (parameters per atom, grid, shift) is axtb_solv::build_model; below, the Born
// radii (born.rs), the smoothed surface (sasa.rs), the Born matrix and kernel gradients
// (kernel.rs) and the energy, potential and gradient of lib.rs:585-730, in the reference's loop
// orders.
struct SolvCache {
  vec brad, dbrad, sasa, hbw, dhbdw, B;  // B row-major nat x nat
  double adet = 0.0, gsasa = 0.0;
  vector<vector<int>> nb;       // SASA neighbours of each atom, ascending
  vector<vector<double>> nbg;   // d sasa_i / d R_j, 3 per neighbour
  vec selfg;                    // d sasa_i / d R_i, 3 per atom
};

struct BornC {
  double g = 0.0, dg = 0.0;
};

// sphere integrals of the GBOBC psi sum (born.rs:20-52); `overlap` sets r1
inline BornC born_nonoverlap(double r, double rho, double r1) {
  const double ap = r + rho, am = r - rho, ab = ap * am, rhab = rho / ab;
  const double lnab = 0.5 * std::log(am / ap) * r1;
  return BornC{rhab + lnab, -2.0 * rhab / ab + (rhab - lnab) * r1 * r1};
}
inline BornC born_overlap(double r, double rho, double rvdw, double& r1) {
  r1 = 1.0 / r;
  const double rr1 = r1, r12 = 0.5 * rr1, r24 = r12 * r12, ap = r + rho, am = r - rho;
  const double rh1 = 1.0 / rvdw, rhr1 = 1.0 / ap, aprh1 = ap * rh1, lnab = std::log(aprh1);
  const double g = rh1 - rhr1 + r12 * (0.5 * am * (rhr1 - rh1 * aprh1) - lnab);
  const double dg = rhr1 * rhr1 * (1.0 - 0.25 * am * rr1 * (1.0 + aprh1 * aprh1)) + rho * r24 * (rhr1 - rh1 * aprh1) +
                    r12 * (rr1 * lnab - rhr1);
  return BornC{g, dg * rr1};
}
// the pair (ii > jj): contributions to psi(ii) and psi(jj) in xtb's four overlap cases, with the
// r1 carried across pairs (born.rs:54-100)
inline void born_pair(double r, double rhoi, double rhoj, double rvdwi, double rvdwj, double& r1, BornC& ci, bool& hi,
                      BornC& cj, bool& hj) {
  const bool ovij = !(r >= rvdwi + rhoj), ovji = !(r >= rhoi + rvdwj);
  hi = hj = false;
  if (!ovij && !ovji) {
    r1 = 1.0 / r;
    if (std::fabs(rhoi - rhoj) < 1.0e-8) {
      ci = cj = born_nonoverlap(r, rhoj, r1);
    } else {
      ci = born_nonoverlap(r, rhoj, r1);
      cj = born_nonoverlap(r, rhoi, r1);
    }
    hi = hj = true;
  } else if (!ovij && ovji) {
    r1 = 1.0 / r;
    ci = born_nonoverlap(r, rhoj, r1);
    hi = true;
    if (r + rhoi > rvdwj) {
      cj = born_overlap(r, rhoi, rvdwj, r1);
      hj = true;
    }
  } else if (ovij && !ovji) {
    if (r + rhoj > rvdwi) {
      ci = born_overlap(r, rhoj, rvdwi, r1);
      hi = true;
    }
    cj = born_nonoverlap(r, rhoi, r1);  // whatever r1 holds (xtb 6.7.1)
    hj = true;
  } else {
    if (r + rhoj > rvdwi) {
      ci = born_overlap(r, rhoj, rvdwi, r1);
      hi = true;
    }
    if (r + rhoi > rvdwj) {
      cj = born_overlap(r, rhoi, rvdwj, r1);
      hj = true;
    }
  }
}

inline void solv_pair_vec(const vec& x, int i, int j, double d[3], double& r2, double& r) {
  for (int k = 0; k < 3; k++) d[k] = x[3 * i + k] - x[3 * j + k];
  r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  r = std::sqrt(r2);
}

// Born radii and dbrad/dpsi (born.rs:112-160)
void solv_born_radii(const axtb_solv::SolvModel& S, const vec& x, SolvCache& c) {
  using namespace axtb_solv;
  const int nat = (int)S.vdwr.size();
  const double lrcut2 = S.lrcut * S.lrcut;
  vec psi(nat, 0.0);
  double r1 = 0.0;
  for (int ii = 0; ii < nat; ii++)
    for (int jj = 0; jj < ii; jj++) {
      double d[3], dr2, r;
      solv_pair_vec(x, ii, jj, d, dr2, r);
      if (!(dr2 < lrcut2)) continue;
      if (S.fresh_r1) r1 = 1.0 / r;
      BornC ci, cj;
      bool hi, hj;
      born_pair(r, S.rho[ii], S.rho[jj], S.vdwr[ii], S.vdwr[jj], r1, ci, hi, cj, hj);
      if (hi) psi[ii] += ci.g;
      if (hj) psi[jj] += cj.g;
    }
  c.brad.assign(nat, 0.0);
  c.dbrad.assign(nat, 0.0);
  for (int i = 0; i < nat; i++) {
    double br = psi[i];
    const double s1 = 1.0 / S.svdw[i], v1 = 1.0 / S.vdwr[i], s2 = 0.5 * S.svdw[i];
    br *= s2;
    double arg2 = br * (OBC_GAMMA * br - OBC_BETA);
    const double arg = br * (OBC_ALPHA + arg2);
    arg2 = 2.0 * arg2 + OBC_ALPHA + OBC_GAMMA * br * br;
    const double th = std::tanh(arg), ch = std::cosh(arg);
    br = 1.0 / (s1 - v1 * th);
    br *= S.born_scale;
    double dpsi = ch * (s1 - v1 * th);
    dpsi = s2 * v1 * arg2 / (dpsi * dpsi);
    dpsi *= S.born_scale;
    c.brad[i] = br;
    c.dbrad[i] = dpsi;
  }
}

// sum_i f_i d psi_i / d R (born.rs:162-207)
void solv_psi_gradient(const axtb_solv::SolvModel& S, const vec& x, const vec& f, vec& g) {
  const int nat = (int)S.vdwr.size();
  const double lrcut2 = S.lrcut * S.lrcut;
  double r1 = 0.0;
  for (int ii = 0; ii < nat; ii++) {
    double gi[3] = {0, 0, 0};
    for (int jj = 0; jj < ii; jj++) {
      double d[3], dr2, r;
      solv_pair_vec(x, ii, jj, d, dr2, r);
      if (!(dr2 < lrcut2)) continue;
      if (S.fresh_r1) r1 = 1.0 / r;
      BornC ci, cj;
      bool hi, hj;
      born_pair(r, S.rho[ii], S.rho[jj], S.vdwr[ii], S.vdwr[jj], r1, ci, hi, cj, hj);
      double s = 0.0;
      if (hi) s += f[ii] * ci.dg;
      if (hj) s += f[jj] * cj.dg;
      if (s != 0.0)
        for (int k = 0; k < 3; k++) {
          gi[k] += s * d[k];
          g[3 * jj + k] -= s * d[k];
        }
    }
    for (int k = 0; k < 3; k++) g[3 * ii + k] += gi[k];
  }
}

// smoothed SASA per atom and its derivatives (sasa.rs:34-80 neighbour lists, 110-225 numsa)
void solv_sasa(const axtb_solv::SolvModel& S, const vec& x, SolvCache& c) {
  const int nat = (int)S.vdwr.size(), np = (int)S.weight.size();
  const double srcut2 = S.srcut * S.srcut;
  c.nb.assign(nat, {});
  for (int i = 0; i < nat; i++)
    for (int j = 0; j < nat; j++) {
      if (j == i) continue;
      double d[3], d2, r;
      solv_pair_vec(x, i, j, d, d2, r);
      const double rmax = S.vdwsa[i] + S.vdwsa[j] + S.w + 1.0e-6;
      if (d2 < srcut2 && d2 < rmax * rmax) c.nb[i].push_back(j);
    }
  c.sasa.assign(nat, 0.0);
  c.selfg.assign(3 * (size_t)nat, 0.0);
  c.nbg.assign(nat, {});
  for (int iat = 0; iat < nat; iat++) {
    const vector<int>& nb = c.nb[iat];
    const int nno = (int)nb.size();
    vec grads(3 * (size_t)nno, 0.0), tj2v(nno);
    double gself[3] = {0, 0, 0}, sasai = 0.0;
    const double rsas = S.vdwsa[iat], wr = S.wrp[iat];
    vector<int> grdi(nno);
    vec grds(3 * (size_t)nno);
    for (int ip = 0; ip < np; ip++) {
      double xp[3];
      for (int k = 0; k < 3; k++) xp[k] = x[3 * iat + k] + rsas * S.grid[3 * ip + k];
      bool buried = false;
      for (int sl = 0; sl < nno; sl++) {
        const int ia = nb[sl];
        double t2 = 0.0;
        for (int k = 0; k < 3; k++) t2 += (xp[k] - x[3 * ia + k]) * (xp[k] - x[3 * ia + k]);
        tj2v[sl] = t2;
        buried |= t2 <= S.trj2a[ia];
      }
      if (buried) continue;
      int nni = 0;
      double sasap = 1.0;
      for (int sl = 0; sl < nno; sl++) {
        const int ia = nb[sl];
        const double tj2 = tj2v[sl];
        if (tj2 < S.trj2b[ia]) {
          double tj[3];
          for (int k = 0; k < 3; k++) tj[k] = xp[k] - x[3 * ia + k];
          const double sqtj = std::sqrt(tj2), uj = sqtj - S.vdwsa[ia], ah3uj2 = S.ah3 * uj * uj;
          double dsasaij = S.ah1 + 3.0 * ah3uj2;
          double sasaij = S.ah0 + (S.ah1 + ah3uj2) * uj;
          if (1.0 < sasaij) sasaij = 1.0;
          if (sasaij < 0.0) sasaij = 0.0;
          if (sasaij < std::numeric_limits<double>::min()) {
            sasap = 0.0;
            break;
          }
          sasap *= sasaij;
          dsasaij /= sasaij * sqtj;
          grdi[nni] = sl;
          for (int k = 0; k < 3; k++) grds[3 * nni + k] = dsasaij * tj[k];
          nni++;
        }
      }
      if (sasap > axtb_solv::TOLSESP) {
        const double wsa = S.weight[ip] * wr * sasap;
        sasai += wsa;
        for (int jj = 0; jj < nni; jj++) {
          const int sl = grdi[jj];
          for (int k = 0; k < 3; k++) {
            const double drjj = wsa * grds[3 * jj + k];
            gself[k] += drjj;
            grads[3 * sl + k] -= drjj;
          }
        }
      }
    }
    c.sasa[iat] = sasai;
    for (int k = 0; k < 3; k++) c.selfg[3 * iat + k] = gself[k];
    c.nbg[iat] = grads;
  }
}

// grad += sum_i w_i d sasa_i / d R (sasa.rs:245-268)
void solv_sasa_contract(const SolvCache& c, const vec& w, vec& g) {
  const int nat = (int)c.sasa.size();
  for (int i = 0; i < nat; i++) {
    const double ci = w[i];
    if (ci == 0.0) continue;
    for (int k = 0; k < 3; k++) g[3 * i + k] += c.selfg[3 * i + k] * ci;
    for (size_t p = 0; p < c.nb[i].size(); p++) {
      const int j = c.nb[i][p];
      for (int k = 0; k < 3; k++) g[3 * j + k] += c.nbg[i][3 * p + k] * ci;
    }
  }
}

// inertia of the atoms weighted by rad^3 (kernel.rs:160-192) and the ALPB shape descriptor
// aDet (kernel.rs:205-209)
void solv_inertia(const vec& x, const vec& rad, double I[3][3], double center[3], double& tot_rad3) {
  const int nat = (int)rad.size();
  tot_rad3 = 0.0;
  for (int k = 0; k < 3; k++) center[k] = 0.0;
  for (int i = 0; i < nat; i++) {
    const double r2 = rad[i] * rad[i], r3 = r2 * rad[i];
    tot_rad3 += r3;
    for (int k = 0; k < 3; k++) center[k] += x[3 * i + k] * r3;
  }
  for (int k = 0; k < 3; k++) center[k] /= tot_rad3;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) I[a][b] = 0.0;
  for (int i = 0; i < nat; i++) {
    const double r2 = rad[i] * rad[i], r3 = r2 * rad[i];
    const double v[3] = {x[3 * i] - center[0], x[3 * i + 1] - center[1], x[3 * i + 2] - center[2]};
    const double vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double diag = r3 * (vv + 0.4 * r2);
    for (int k = 0; k < 3; k++) I[k][k] += diag;
    for (int col = 0; col < 3; col++)
      for (int row = 0; row < 3; row++) I[row][col] -= r3 * v[col] * v[row];
  }
}
inline double solv_det3(const double a[3][3]) {
  return a[0][0] * a[1][1] * a[2][2] - a[0][0] * a[1][2] * a[2][1] - a[0][1] * a[1][0] * a[2][2] +
         a[0][1] * a[1][2] * a[2][0] + a[0][2] * a[1][0] * a[2][1] - a[0][2] * a[1][1] * a[2][0];
}
double solv_adet(const vec& x, const vec& rad) {
  double I[3][3], c[3], t;
  solv_inertia(x, rad, I, c, t);
  return std::sqrt(std::pow(solv_det3(I), 1.0 / 3.0) / (0.4 * t));
}
// gradient of keps alpbet / (2 aDet) (sum q)^2 (kernel.rs:211-238)
void solv_adet_deriv(const vec& x, const vec& rad, double keps_alpbet, const vec& q, vec& g) {
  double qt = 0.0;
  for (double v : q) qt += v;
  double i[3][3], center[3], t;
  solv_inertia(x, rad, i, center, t);
  const double adet = std::sqrt(std::pow(solv_det3(i), 1.0 / 3.0) / (0.4 * t));
  const double t2 = t * t, a2 = adet * adet;
  const double fa = 250.0 / (48.0 * (t2 * t) * ((a2 * a2) * adet));
  const double fb = -0.5 * keps_alpbet * (qt * qt) / a2;
  auto f = [&](double v) { return v * fa * fb; };
  const double xy = i[0][1] * i[2][2] - i[0][2] * i[1][2], xz = i[0][2] * i[1][1] - i[0][1] * i[2][1];
  const double yz = i[0][0] * i[1][2] - i[0][1] * i[0][2];
  const double xx = i[0][0] * (i[1][1] + i[2][2]) - i[0][1] * i[0][1] - i[0][2] * i[0][2];
  const double yy = i[1][1] * (i[0][0] + i[2][2]) - i[0][1] * i[0][1] - i[1][2] * i[1][2];
  const double zz = i[2][2] * (i[0][0] + i[1][1]) - i[0][2] * i[0][2] - i[1][2] * i[1][2];
  const double d[3][3] = {{f(xx), f(xy), f(xz)}, {f(xy), f(yy), f(yz)}, {f(xz), f(yz), f(zz)}};
  for (size_t a = 0; a < rad.size(); a++) {
    const double r3 = rad[a] * rad[a] * rad[a];
    const double v[3] = {x[3 * a] - center[0], x[3 * a + 1] - center[1], x[3 * a + 2] - center[2]};
    for (int k = 0; k < 3; k++) g[3 * a + k] += r3 * (d[k][0] * v[0] + d[k][1] * v[1] + d[k][2] * v[2]);
  }
}

// geometry-dependent intermediates (lib.rs:555-613)
void solv_update(const axtb_solv::SolvModel& S, const vec& x, SolvCache& c) {
  using namespace axtb_solv;
  const int nat = (int)S.vdwr.size();
  solv_born_radii(S, x, c);
  solv_sasa(S, x, c);
  c.gsasa = 0.0;
  for (int i = 0; i < nat; i++) c.gsasa += c.sasa[i] * S.gamsasa[i];
  c.hbw.assign(nat, 0.0);
  c.dhbdw.assign(nat, 0.0);
  if (S.lhb)
    for (int i = 0; i < nat; i++) {
      const double smaxd = 1.0 / (S.vdwsa[i] * S.vdwsa[i]);
      c.hbw[i] = S.hbmag[i] * c.sasa[i] * smaxd;
      c.dhbdw[i] = S.hbmag[i] * smaxd;
    }
  c.adet = S.alpbet > 0.0 ? solv_adet(x, S.vdwr) : 0.0;
  // Born matrix: P16 (ALPB) or Still (GBSA) kernel, self terms keps / a_i (kernel.rs:28-70)
  c.B.assign((size_t)nat * nat, 0.0);
  const double z16 = ZETA_P16 / 16.0;
  for (int i = 0; i < nat; i++)
    for (int j = 0; j < i; j++) {
      double d[3], r2, r;
      solv_pair_vec(x, i, j, d, r2, r);
      double v;
      if (S.alpb) {
        const double ab = std::sqrt(c.brad[i] * c.brad[j]);
        double arg = ab / (ab + z16 * r);
        arg *= arg;
        arg *= arg;
        arg *= arg;
        arg *= arg;
        v = S.keps * (1.0 / (r + ab * arg));
      } else {
        const double rr2 = r * r, aa = c.brad[i] * c.brad[j], dd = 0.25 * rr2 / aa;
        v = S.keps * (1.0 / std::sqrt(rr2 + aa * std::exp(-dd)));
      }
      c.B[(size_t)i * nat + j] = c.B[(size_t)j * nat + i] = v;
    }
  for (int i = 0; i < nat; i++) c.B[(size_t)i * nat + i] = S.keps * (1.0 / c.brad[i]);
  if (S.lhb)
    for (int i = 0; i < nat; i++) c.B[(size_t)i * nat + i] += 2.0 * c.hbw[i];
  if (S.alpbet > 0.0) {
    const double shift = S.keps * S.alpbet / c.adet;
    for (double& b : c.B) b += shift;
  }
}

// (B q)_i in the reference's dot order (lib.rs:760-779: four interleaved partial sums)
inline double solv_row_dot(const double* row, const double* q, int n) {
  double acc[4] = {0, 0, 0, 0};
  int k = 0;
  for (; k + 4 <= n; k += 4)
    for (int t = 0; t < 4; t++) acc[t] += row[k + t] * q[k + t];
  double s = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  for (; k < n; k++) s += row[k] * q[k];
  return s;
}

// explicit gradient of 1/2 q^T B q at fixed charges (lib.rs:688-709, kernel.rs:75-157)
void solv_gradient(const axtb_solv::SolvModel& S, const vec& x, const SolvCache& c, const vec& q, vec& g) {
  using namespace axtb_solv;
  const int nat = (int)S.vdwr.size();
  vec dedbr(nat, 0.0);
  const double keps = S.keps;
  for (int iat = 0; iat < nat; iat++) {
    double gi[3] = {0, 0, 0}, di = 0.0;
    for (int jat = 0; jat < iat; jat++) {
      double v[3], r2, r;
      solv_pair_vec(x, iat, jat, v, r2, r);
      const double qq = q[iat] * q[jat];
      if (S.alpb) {
        const double ab = std::sqrt(c.brad[iat] * c.brad[jat]);
        const double arg1 = ab / (ab + ZETA_P16 / 16.0 * r);
        double arg16 = arg1 * arg1;
        arg16 *= arg16;
        arg16 *= arg16;
        arg16 *= arg16;
        const double fgb = r + ab * arg16, dfgb = 1.0 / fgb, dfgb2 = dfgb * dfgb;
        const double ap = (1.0 - ZETA_P16 * arg1 * arg16) * dfgb2;
        for (int k = 0; k < 3; k++) {
          const double dg = ap * v[k] * keps / r * qq;
          gi[k] -= dg;
          g[3 * jat + k] += dg;
        }
        const double bp = -0.5 * (r * ZETA_P16 / ab * arg1 + 1.0) / ab * arg16 * dfgb2;
        di += c.brad[jat] * bp * keps * qq;
        dedbr[jat] += c.brad[iat] * bp * keps * qq;
      } else {
        const double rr2 = r * r, aa = c.brad[iat] * c.brad[jat], dd = 0.25 * rr2 / aa, expd = std::exp(-dd);
        const double fgb2 = rr2 + aa * expd, dfgb2 = 1.0 / fgb2, dfgb = std::sqrt(dfgb2);
        const double dfgb3 = dfgb2 * dfgb * keps, ap = (1.0 - 0.25 * expd) * dfgb3;
        for (int k = 0; k < 3; k++) {
          const double dr = ap * v[k];
          gi[k] -= dr * qq;
          g[3 * jat + k] += dr * qq;
        }
        const double bp = -0.5 * expd * (1.0 + dd) * dfgb3;
        di += c.brad[jat] * bp * qq;
        dedbr[jat] += c.brad[iat] * bp * qq;
      }
    }
    for (int k = 0; k < 3; k++) g[3 * iat + k] += gi[k];
    dedbr[iat] += di;
  }
  for (int i = 0; i < nat; i++) {
    const double bp = 1.0 / c.brad[i], qq = q[i] * bp;
    dedbr[i] += -0.5 * keps * qq * bp * q[i];
  }
  vec work(nat);
  for (int i = 0; i < nat; i++) work[i] = dedbr[i] * c.dbrad[i];
  solv_psi_gradient(S, x, work, g);
  if (S.lhb) {
    for (int i = 0; i < nat; i++) work[i] = c.dhbdw[i] * q[i] * q[i];
    solv_sasa_contract(c, work, g);
  }
  if (S.alpbet > 0.0) solv_adet_deriv(x, S.vdwr, S.keps * S.alpbet, q, g);
}

// energy parts as xtb prints them: gsolv, gelec, gsasa, ghb, gshift (lib.rs:711-735)
void solv_parts(const axtb_solv::SolvModel& S, const SolvCache& c, const vec& q, double out[5]) {
  const int nat = (int)q.size();
  double ghb = 0.0, epol = 0.0;
  if (S.lhb)
    for (int i = 0; i < nat; i++) ghb += c.hbw[i] * (q[i] * q[i]);
  for (int i = 0; i < nat; i++) epol += q[i] * (0.5 * solv_row_dot(&c.B[(size_t)i * nat], q.data(), nat));
  const double gelec = epol - ghb;
  out[0] = gelec + c.gsasa + ghb + S.gshift;
  out[1] = gelec;
  out[2] = c.gsasa;
  out[3] = ghb;
  out[4] = S.gshift;
}

// ------------------------------------------------------------------ the calculation (xtb.rs:437-861)
struct Options {
  double acc = 1.0, etemp = 300.0, damping = 0.4;
  int max_iter = 250, guess = 0 /*0 EEQ, 1 SAD*/, want_grad = 1, dispersion = 1;
  int mixer = MIX_AUTO;  // the axtb conventions' plan (conventions.rs:260)
  // conventions: 0 axtb, 3 tblite 0.8.0 (conventions.rs:73-84)
  int conventions = 0;
  int uhf = -1;  // unpaired electrons; -1: the lowest spin of the electron count
  // a given start (ScfGuess, xtb.rs:204-208): shell charges [nsh] (null: none), dipoles [3N],
  // quadrupoles [6N] (null: zero)
  const double *gq = nullptr, *gdp = nullptr, *gqp = nullptr;
  // implicit solvation (SolvationSpec, solvation.rs:14-22): model 0 none / 1 ALPB / 2 GBSA, the
  // solvent's name, reference state; the axtb conventions' corrected model (conventions.rs:256)
  int solv_model = 0, solv_state = 0;
  const char* solvent = nullptr;
};

struct Result {
  double energy = 0;
  double terms[9] = {0};  // repulsion dispersion electronic isotropic_es third_order anisotropic entropy band solvation
  vec grad, qat, qsh, dp, qp, emo, focc;
  int iterations = 0, nao = 0, nsh = 0, converged = 0;
  double efermi = 0, residual = 0, energy_change = 0;
  // the attempts of the mixer plan (iterations, converged) and the mixer of the result:
  // 0 broyden, 1 diis, 2 diis-fallback (MixerPlan::used, scf/mixing.rs:150-157)
  int nattempts = 0, attempt_iters[2] = {0, 0}, attempt_conv[2] = {0, 0}, mixer_used = 0;
  double dipole[3] = {0, 0, 0}, gap = 0;
};

struct DebugOut {
  double *S, *H0, *D, *Q, *cn, *qstart, *vsh, *vat, *vdp, *vqp, *F, *w, *P, *qsh;
};
DebugOut debug_out;
bool debug_on = false;
inline void dbg_copy(double* dst, const double* src, size_t n) {
  if (debug_on && dst) std::memcpy(dst, src, n * sizeof(double));
}

// P = C diag(f) C^T over the orbitals with f > 1e-16 (density.rs:65-87: the lower triangle, mirrored);
// with eps: W = C diag(f eps) C^T (xtb.rs:688-696). Tmp: n x n work.
void weighted_outer(const Mat& C, const vec& focc, const vec* eps, Mat& Tmp, Mat& P) {
  const int n = C.n;
  vector<int> kept;
  for (int k = 0; k < n; k++)
    if (focc[k] > 1e-16) kept.push_back(k);
  const int kk = (int)kept.size();
  Mat Ck(n, std::max(kk, 1));
#pragma omp parallel for schedule(static)
  for (int q = 0; q < kk; q++) {
    const int k = kept[q];
    const double f = eps ? focc[k] * (*eps)[k] : focc[k];
    for (int i = 0; i < n; i++) {
      Ck(i, q) = C(i, k);
      Tmp(i, q) = C(i, k) * f;
    }
  }
  mm(false, true, n, n, kk, 1.0, Ck.a.data(), n, Tmp.a.data(), n, 0.0, P.a.data(), n);
  mirror_lower(P);
}

const double SQRT2 = 1.4142135623730951;
const bool QOFF[6] = {false, true, false, true, true, false};

Result singlepoint(int nat, const int* z, const double* xyz, double charge, const Options& opt,
                   Result* report = nullptr) {
  if (!(opt.acc > 0.0 && std::isfinite(opt.acc))) throw Error(1, "accuracy must be positive");
  if (!(opt.etemp >= 0.0 && std::isfinite(opt.etemp))) throw Error(1, "etemp must be >= 0");
  if (!(opt.damping > 0.0 && opt.damping <= 1.0)) throw Error(1, "damping must be in (0, 1]");
  if (opt.max_iter <= 0) throw Error(1, "max_iter must be >= 1");
  const bool tbl = opt.conventions == 3;
  if (opt.conventions != 0 && !tbl) throw Error(1, "the oracle runs the axtb and tblite conventions");
  Model m = build_model(nat, z, xyz, charge, opt.acc, tbl);
  int nsh = m.nsh, n = m.nao;
  // electrons (xtb.rs:1178-1212), closed shell / lowest spin
  double nref = 0.0;
  for (int s = 0; s < nsh; s++) nref += m.refocc[s];
  if (std::fmod(charge, 1.0) != 0.0) throw Error(1, "the charge must be an integer");
  double nocc = nref - charge;
  long nel_int = std::lround(nocc);
  // unpaired electrons: given (xtb.rs:1178-1212, checked against the electron count's parity),
  // else the lowest spin
  if (opt.uhf >= 0 && (opt.uhf % 2) != (int)(((nel_int % 2) + 2) % 2))
    throw Error(1, "electron count and unpaired electrons are not compatible");
  double nuhf = opt.uhf >= 0 ? (double)opt.uhf : (double)(((nel_int % 2) + 2) % 2);
  if (nocc < 0.0) throw Error(1, "negative electron count");
  if (nuhf > nocc) throw Error(1, "too many unpaired electrons");
  double nalpha = (nocc + nuhf) / 2.0, nbeta = (nocc - nuhf) / 2.0;
  if (nalpha > n + 1e-9) throw Error(1, "too many electrons for the basis");
  double kt = opt.etemp * (tbl ? KB_XTB_TBLITE : KB_HARTREE);

  vec g(3 * nat, 0.0), dedcn(nat, 0.0);
  vec energies(nat, 0.0), erep(nat, 0.0);
  // geometry (xtb.rs:275-292)
  vec cn = gfn2_cn(m);
  vec se(nsh), dsedcn(nsh);
  for (int s = 0; s < nsh; s++) {
    se[s] = m.se0[s] - m.kcn[s] * cn[m.sh2at[s]];
    dsedcn[s] = -m.kcn[s];
  }
  Coulomb cou = build_coulomb(m, cn);
  D4Options d4o;
  if (tbl) {  // mctc-lib's radii and D4's cutoffs (RealspaceCutoff::TBLITE_GFN2, disp.rs:47)
    d4o.rcov_mode = 0;
    d4o.cn_cut = 30.0;
    d4o.disp2 = 50.0;
    d4o.disp3 = 25.0;
    d4o.width2 = d4o.width3 = 0.05;
  }
  D4 d4 = build_d4(nat, z, xyz, d4o);
  Result res;
  res.terms[0] = repulsion(m, erep, opt.want_grad ? &g : nullptr);
  vec e3b(nat, 0.0);
  if (opt.dispersion) d4_three_body(d4, e3b, opt.want_grad ? &g : nullptr);
  double e3sum = 0;
  for (int i = 0; i < nat; i++) {
    energies[i] += e3b[i];
    e3sum += e3b[i];
  }
  for (int i = 0; i < nat; i++) energies[i] += erep[i];
  // implicit solvation: the surface and shift energies are charge independent (lib.rs:615-636)
  axtb_solv::SolvModel solv;
  SolvCache sc;
  const bool solvated = opt.solv_model != 0;
  double esolv_fixed = 0.0;
  if (solvated) {
    std::string err;
    if (!axtb_solv::build_model(nat, z, opt.solv_model, opt.solvent, opt.solv_state, axtb_solv::CONV_CORRECTED,
                                axtb_solv::DEFAULT_NANG, axtb_solv::DEFAULT_TEMPERATURE, solv, err))
      throw Error(1, "solvation: " + err);
    solv_update(solv, m.xyz, sc);
    const double shift = solv.gshift / nat;
    for (int i = 0; i < nat; i++) {
      const double e = sc.sasa[i] * solv.gamsasa[i] + shift;
      energies[i] += e;
      esolv_fixed += e;
    }
  }

  Ints I = build_integrals(m, se);
  {
    size_t nn = (size_t)m.nao * m.nao;
    dbg_copy(debug_out.S, I.S.a.data(), nn);
    dbg_copy(debug_out.H0, I.H0.a.data(), nn);
    for (int k = 0; k < 3; k++) dbg_copy(debug_out.D ? debug_out.D + k * nn : nullptr, I.D[k].a.data(), nn);
    for (int k = 0; k < 6; k++) dbg_copy(debug_out.Q ? debug_out.Q + k * nn : nullptr, I.Q[k].a.data(), nn);
    dbg_copy(debug_out.cn, cn.data(), nat);
  }
  // Cholesky S = L L^T (eigen.rs:72-113)
  // (right-looking: column j scaled, then the trailing columns updated, each on contiguous memory)
  Mat L = I.S;
  for (int j = 0; j < n; j++) {
    const double s = L(j, j);
    if (!(s > 0.0)) throw Error(2, "overlap matrix is not positive definite");
    const double ljj = std::sqrt(s);
    double* lj = &L.a[(size_t)j * n];
    lj[j] = ljj;
    for (int i = j + 1; i < n; i++) lj[i] /= ljj;
#pragma omp parallel for schedule(static, 8)
    for (int c = j + 1; c < n; c++) {
      double* lc = &L.a[(size_t)c * n];
      const double f = lj[c];
      if (f != 0.0)
        for (int i = c; i < n; i++) lc[i] -= lj[i] * f;
    }
  }
  for (int j = 0; j < n; j++)
    for (int i = 0; i < j; i++) L(i, j) = 0.0;
  // Linv (lower): L X = I by forward substitution, eight columns of X per pass over L
  Mat Li(n, n);
#pragma omp parallel for schedule(dynamic, 1)
  for (int j0 = 0; j0 < n; j0 += 8) {
    const int j1 = std::min(n, j0 + 8);
    for (int j = j0; j < j1; j++) Li(j, j) = 1.0;
    for (int k = j0; k < n; k++) {
      const double* lk = &L.a[(size_t)k * n];
      for (int j = j0; j < j1 && j <= k; j++) {
        double* x = &Li.a[(size_t)j * n];
        const double xk = (x[k] /= lk[k]);
        if (xk != 0.0)
          for (int i = k + 1; i < n; i++) x[i] -= lk[i] * xk;
      }
    }
  }

  // SCC vector layout (scf/variables.rs): [qsh | dp (3N) | qp (6N, off-diagonals x sqrt2)]
  int ndim = nsh + 9 * nat;
  auto pack = [&](const Moments& s, vec& out) {
    for (int i = 0; i < nsh; i++) out[i] = s.qsh[i];
    for (int a = 0; a < nat; a++) {
      for (int c = 0; c < 3; c++) out[nsh + 3 * a + c] = s.dp[3 * a + c];
      for (int c = 0; c < 6; c++) out[nsh + 3 * nat + 6 * a + c] = s.qp[6 * a + c] * (QOFF[c] ? SQRT2 : 1.0);
    }
  };
  auto qat_of = [&](Moments& s) {
    s.qat.assign(nat, 0.0);
    for (int i = 0; i < nsh; i++) s.qat[m.sh2at[i]] += s.qsh[i];
  };
  auto unpack = [&](const vec& v, Moments& s) {
    for (int i = 0; i < nsh; i++) s.qsh[i] = v[i];
    for (int a = 0; a < nat; a++) {
      for (int c = 0; c < 3; c++) s.dp[3 * a + c] = v[nsh + 3 * a + c];
      for (int c = 0; c < 6; c++) s.qp[6 * a + c] = v[nsh + 3 * nat + 6 * a + c] / (QOFF[c] ? SQRT2 : 1.0);
    }
    qat_of(s);
  };
  Moments start;
  start.qsh.assign(nsh, 0.0);
  start.dp.assign(3 * nat, 0.0);
  start.qp.assign(6 * nat, 0.0);
  if (opt.gq) {
    // the given moments (xtb.rs:522-528)
    for (int s = 0; s < nsh; s++) start.qsh[s] = opt.gq[s];
    if (opt.gdp) start.dp.assign(opt.gdp, opt.gdp + 3 * nat);
    if (opt.gqp) start.qp.assign(opt.gqp, opt.gqp + 6 * nat);
  } else if (opt.guess == 0 && !tbl) {  // tblite starts from zero charges (conventions.rs:335)
    try {
      vec q = eeq_charges(nat, z, xyz, charge);
      for (int iat = 0; iat < nat; iat++) {
        int is = m.ish_at[iat], ns = m.nsh_at[iat];
        double tot = 0.0;
        for (int s = 0; s < ns; s++) tot += m.refocc[is + s];
        for (int s = 0; s < ns; s++) start.qsh[is + s] = tot > 0.0 ? q[iat] * m.refocc[is + s] / tot : 0.0;
      }
    } catch (Error&) {
    }
  }
  qat_of(start);
  dbg_copy(debug_out.qstart, start.qsh.data(), nsh);

  Moments st = start, out = start;
  Pots pot;
  Mat F(n, n), C(n, n), P(n, n), W(n, n), Tmp(n, n);
  vec emo(n), occa(n), occb(n), focc(n), vao(n);
  double ts = 0.0, elast = 0.0, efermi = 0.0;
  vec eelec(nat);
  vec packed(ndim);
  bool converged = false;
  int iter = 0;
  double res_norm = 0, de_last = 0;
  // one SCF per mixer of the plan, each from the start (xtb.rs:590-646): auto = Broyden, then DIIS
  // the conventions' default mixer (xtb.rs:516): axtb Broyden then DIIS, tblite Broyden
  const int mixer = opt.mixer >= 0 ? opt.mixer : (tbl ? MIX_BROYDEN : MIX_AUTO);
  const int plan[2] = {mixer == MIX_DIIS ? MIX_DIIS : MIX_BROYDEN, MIX_DIIS};
  const int nplan = mixer == MIX_AUTO ? 2 : 1;
  int total_iters = 0;
  for (int ia = 0; ia < nplan && !converged; ia++) {
  const bool last_attempt = ia + 1 == nplan;
  VecMixer mix(plan[ia], ndim, opt.max_iter, opt.damping);
  st = start;
  elast = 0.0;  // run_scf's first iteration compares with energy 0 (scf/driver.rs:191)
  std::fill(eelec.begin(), eelec.end(), 0.0);
  int stop = 0;  // 0 iteration limit or converged, 1 mixer failed, 2 Fermi level not found
  int iters = 0;
  for (iter = 1; iter <= opt.max_iter; iter++) {
    if (iter > 1) {
      if (!mix.next()) {
        stop = 1;
        break;
      }
      unpack(mix.q_in(), st);
    }
    iters = iter;
    // potentials (xtb.rs:316-340)
    pot.vsh.assign(nsh, 0.0);
    pot.vat.assign(nat, 0.0);
    pot.vdp.assign(3 * nat, 0.0);
    pot.vqp.assign(6 * nat, 0.0);
    for (int a = 0; a < nsh; a++) {
      double s = 0.0;
      for (int b = 0; b < nsh; b++) s += cou.gam[(size_t)a * nsh + b] * st.qsh[b];
      pot.vsh[a] += s;
    }
    for (int a = 0; a < nsh; a++) pot.vsh[a] += st.qsh[a] * st.qsh[a] * cou.gamma3[a];
    aes_potentials(m, cou, st, pot);
    if (opt.dispersion) d4_energy_potential(d4, st.qat, nullptr, &pot.vat);
    if (solvated)  // B q of the input charges (lib.rs:653-661)
      for (int i = 0; i < nat; i++) pot.vat[i] += solv_row_dot(&sc.B[(size_t)i * nat], st.qat.data(), nat);
    for (int mu = 0; mu < n; mu++) vao[mu] = pot.vsh[m.ao2sh[mu]] + pot.vat[m.ao2at[mu]];
    // Fock (scf/fock.rs:22-104)
#pragma omp parallel for schedule(static)
    for (int nu = 0; nu < n; nu++) {
      int at = m.ao2at[nu];
      for (int mu = 0; mu < n; mu++) {
        double b = I.S(mu, nu) * vao[nu];
        for (int k = 0; k < 3; k++) b += I.D[k](mu, nu) * pot.vdp[3 * at + k];
        for (int k = 0; k < 6; k++) b += I.Q[k](mu, nu) * pot.vqp[6 * at + k];
        Tmp(mu, nu) = b;
      }
    }
    bool finite = true;
#pragma omp parallel for schedule(static) reduction(&& : finite)
    for (int nu = 0; nu < n; nu++)
      for (int mu = 0; mu < n; mu++) {
        F(mu, nu) = I.H0(mu, nu) - 0.5 * (Tmp(mu, nu) + Tmp(nu, mu));
        finite = finite && std::isfinite(F(mu, nu));
      }
    if (!finite) throw Error(2, "non-finite Fock matrix");
    pack(st, mix.q_in());
    if (iter == 1) {
      dbg_copy(debug_out.vsh, pot.vsh.data(), nsh);
      dbg_copy(debug_out.vat, pot.vat.data(), nat);
      dbg_copy(debug_out.vdp, pot.vdp.data(), 3 * nat);
      dbg_copy(debug_out.vqp, pot.vqp.data(), 6 * nat);
      dbg_copy(debug_out.F, F.a.data(), (size_t)n * n);
    }
    // A = Linv F Linv^T, eig, C = Linv^T U
    mm(false, false, n, n, n, 1.0, Li.a.data(), n, F.a.data(), n, 0.0, Tmp.a.data(), n);  // Linv F
    vec A((size_t)n * n);
    mm(false, true, n, n, n, 1.0, Tmp.a.data(), n, Li.a.data(), n, 0.0, A.data(), n);  // (Linv F) Linv^T
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < j; i++) {
        double v = 0.5 * (A[(size_t)j * n + i] + A[(size_t)i * n + j]);
        A[(size_t)j * n + i] = A[(size_t)i * n + j] = v;
      }
    eigh(n, A, emo);
    mm(true, false, n, n, n, 1.0, Li.a.data(), n, A.data(), n, 0.0, C.a.data(), n);  // Linv^T U
    // Fermi filling (xtb.rs:997-1031)
    if (kt > 0.0) {
      efermi = fill_tblite(nalpha, kt, emo, occa);
      fill_tblite(nbeta, kt, emo, occb);
    } else {
      aufbau(nalpha, occa);
      aufbau(nbeta, occb);
    }
    double pa = 0, pb = 0;
    for (int i = 0; i < n; i++) {
      pa += occa[i];
      pb += occb[i];
    }
    if (std::fabs(pa - nalpha) > 1e-6 || std::fabs(pb - nbeta) > 1e-6) {
      stop = 2;  // ends this attempt; the last attempt reports it (xtb.rs:616-621)
      break;
    }
    ts = entropy_tblite(occa, kt) + entropy_tblite(occb, kt);
    for (int i = 0; i < n; i++) focc[i] = occa[i] + occb[i];
    // density P = C f C^T (xtb.rs:1033-1038)
    weighted_outer(C, focc, nullptr, Tmp, P);
    // Mulliken (scf/mulliken.rs:19-59) and the band energy per atom
    out.qsh = m.refocc;
    out.dp.assign(3 * nat, 0.0);
    out.qp.assign(6 * nat, 0.0);
    vec eband(nat, 0.0);
    double band = 0.0;
    vec cols((size_t)11 * n);
#pragma omp parallel for schedule(static)
    for (int nu = 0; nu < n; nu++) {
      double r[11] = {0};
      for (int mu = 0; mu < n; mu++) {
        double p = P(mu, nu);
        r[0] += p * I.S(mu, nu);
        for (int k = 0; k < 3; k++) r[1 + k] += p * I.D[k](mu, nu);
        for (int k = 0; k < 6; k++) r[4 + k] += p * I.Q[k](mu, nu);
        r[10] += p * I.H0(mu, nu);
      }
      for (int k = 0; k < 11; k++) cols[(size_t)11 * nu + k] = r[k];
    }
    for (int nu = 0; nu < n; nu++) {
      const double* r = &cols[(size_t)11 * nu];
      int iat = m.ao2at[nu];
      out.qsh[m.ao2sh[nu]] -= r[0];
      for (int k = 0; k < 3; k++) out.dp[3 * iat + k] -= r[1 + k];
      for (int k = 0; k < 6; k++) out.qp[6 * iat + k] -= r[4 + k];
      eband[iat] += r[10];
      band += r[10];
    }
    qat_of(out);
    if (iter == 1) {
      dbg_copy(debug_out.w, emo.data(), n);
      dbg_copy(debug_out.P, P.a.data(), (size_t)n * n);
      dbg_copy(debug_out.qsh, out.qsh.data(), nsh);
    }
    pack(out, packed);
    mix.set_residual(packed);
    // energy of the output moments (xtb.rs:1058-1146)
    std::fill(eelec.begin(), eelec.end(), 0.0);
    double es2 = 0.0;
    for (int a = 0; a < nsh; a++) {
      double v = 0.0;
      for (int b = 0; b < nsh; b++) v += cou.gam[(size_t)a * nsh + b] * out.qsh[b];
      es2 += out.qsh[a] * v;
      eelec[m.sh2at[a]] += 0.5 * out.qsh[a] * v;
    }
    es2 *= 0.5;
    vec e3(nat, 0.0), eaes(nat, 0.0), ed(nat, 0.0);
    for (int a = 0; a < nsh; a++) e3[m.sh2at[a]] += std::pow(out.qsh[a], 3) * cou.gamma3[a] / 3.0;
    aes_energies(m, cou, out, eaes);
    if (opt.dispersion) d4_energy_potential(d4, out.qat, &ed, nullptr);
    vec esq(nat, 0.0);
    double ssolv = 0.0;
    if (solvated)  // 1/2 q_i (B q)_i of the output charges (lib.rs:640-650)
      for (int i = 0; i < nat; i++) {
        esq[i] = out.qat[i] * (0.5 * solv_row_dot(&sc.B[(size_t)i * nat], out.qat.data(), nat));
        ssolv += esq[i];
      }
    double se3 = 0, saes = 0, sd = 0;
    for (int i = 0; i < nat; i++) {
      se3 += e3[i];
      saes += eaes[i];
      sd += ed[i];
      eelec[i] += e3[i] + eaes[i] + ed[i] + esq[i] + ts / nat + eband[i];
    }
    double energy = 0.0;
    for (int i = 0; i < nat; i++) energy += eelec[i];
    if (!std::isfinite(energy)) throw Error(2, "non-finite energy");
    res_norm = 0.0;
    for (double x : mix.dq()) res_norm += x * x;
    res_norm = std::sqrt(res_norm / ndim);
    de_last = energy - elast;
    elast = energy;
    res.terms[2] = 0;
    for (int i = 0; i < nat; i++) res.terms[2] += eelec[i];
    res.terms[1] = e3sum + sd;
    res.terms[3] = es2;
    res.terms[4] = se3;
    res.terms[5] = saes;
    res.terms[6] = ts;
    res.terms[7] = band;
    res.terms[8] = esolv_fixed + ssolv;
    // convergence (scf/convergence.rs:99-121)
    if (std::fabs(de_last) < 1e-6 * opt.acc && res_norm < 2e-5 * opt.acc) {
      converged = true;
      break;
    }
  }
  res.attempt_iters[ia] = iters;
  res.attempt_conv[ia] = converged ? 1 : 0;
  res.nattempts = ia + 1;
  total_iters += iters;
  iter = iters;
  res.mixer_used = mixer == MIX_AUTO && res.nattempts > 1 ? 2 : plan[res.nattempts - 1];
  res.iterations = total_iters;
  if (report) {
    report->iterations = res.iterations;
    report->mixer_used = res.mixer_used;
    report->nattempts = res.nattempts;
    for (int k = 0; k < 2; k++) {
      report->attempt_iters[k] = res.attempt_iters[k];
      report->attempt_conv[k] = res.attempt_conv[k];
    }
  }
  if (stop == 1 && last_attempt)
    throw Error(2, plan[ia] == MIX_BROYDEN ? "Broyden mixing failed" : "DIIS mixing failed");
  if (stop == 2 && last_attempt) throw Error(2, "Fermi level not found");
  }
  if (!converged) throw Error(3, "SCF not converged in " + std::to_string(total_iters) + " iterations");
  res.iterations = iter;
  res.converged = 1;
  res.residual = res_norm;
  res.energy_change = de_last;
  double etot = 0.0;
  for (int i = 0; i < nat; i++) etot += energies[i] + eelec[i];
  res.energy = etot;
  res.nao = n;
  res.nsh = nsh;
  res.qsh = out.qsh;
  res.qat = out.qat;
  res.dp = out.dp;
  res.qp = out.qp;
  res.emo = emo;
  res.focc = focc;
  res.efermi = efermi;
  // molecular dipole about the origin (xtb.rs:746-751) and the HOMO-LUMO gap above the alpha HOMO
  // of the one set of orbitals (properties.rs:204-221)
  for (int i = 0; i < nat; i++)
    for (int k = 0; k < 3; k++) res.dipole[k] += out.qat[i] * m.xyz[3 * i + k] + out.dp[3 * i + k];
  {
    const double nel = std::max(nalpha, nbeta), fl = std::floor(nel);
    const double kk = nel - fl > 0.5 ? fl + 1.0 : fl;
    res.gap = (kk >= 1.0 && (int)kk < n) ? emo[(int)kk] - emo[(int)kk - 1] : std::nan("");
  }
  if (opt.want_grad) {
    // W = C diag(f eps) C^T (xtb.rs:688-696)
    weighted_outer(C, focc, &emo, Tmp, W);
    // add_scc_gradient (xtb.rs:346-387), output moments, last input potentials
    es2_gradient(m, cou, out.qsh, g);
    aes_gradient(m, cou, out, g, dedcn);
    if (opt.dispersion) d4_gradient(d4, out.qat, g);
    if (solvated) {
      // the charge-dependent part at the output charges, and the surface term (lib.rs:615-636,688-709)
      solv_gradient(solv, m.xyz, sc, out.qat, g);
      solv_sasa_contract(sc, solv.gamsasa, g);
    }
    // H0/S/multipole gradient (hamiltonian.rs:137-223)
    // atom pairs by the first atom over the threads, each with its own sums, added in thread order
    const int nt = std::max(1, omp_get_max_threads());
    vector<vec> tg(nt, vec(3 * nat, 0.0)), tc(nt, vec(nat, 0.0));
#pragma omp parallel
    {
    MpGrad mg;
    vec& g = tg[omp_get_thread_num()];
    vec& dedcn = tc[omp_get_thread_num()];
#pragma omp for schedule(static, 1)
    for (int iat = 0; iat < nat; iat++) {
      for (int jat = 0; jat < iat; jat++) {
        double vec3[3];
        for (int k = 0; k < 3; k++) vec3[k] = m.xyz[3 * iat + k] - m.xyz[3 * jat + k];
        double r2 = dot3(vec3, vec3);
        if (r2 > m.pair_cut2) continue;
        if (pair_screened(m.min_alpha_at[iat], m.min_alpha_at[jat], r2, m.intcut)) continue;
        if (r2 < 2.220446049250313e-16) continue;
        double rr = std::sqrt(std::sqrt(r2) / (m.rad[jat] + m.rad[iat]));
        double dg[3] = {0, 0, 0}, dcni = 0, dcnj = 0;
        for (int ish = 0; ish < m.nsh_at[iat]; ish++) {
          int si = m.ish_at[iat] + ish;
          const Cgto& ci = m.shells[iat][ish];
          for (int jsh = 0; jsh < m.nsh_at[jat]; jsh++) {
            int sj = m.ish_at[jat] + jsh;
            const Cgto& cj = m.shells[jat][jsh];
            if (pair_screened(m.min_alpha_sh[si], m.min_alpha_sh[sj], r2, m.intcut)) continue;
            multipole_grad(cj, ci, vec3, m.intcut, mg);
            double shpi = 1.0 + m.shpoly[si] * rr, shpj = 1.0 + m.shpoly[sj] * rr;
            double shp = shpi * shpj;
            double dshp = (shpi * m.shpoly[sj] + shpj * m.shpoly[si]) * 0.5 * rr / r2;
            double dsv[3];
            for (int k = 0; k < 3; k++) dsv[k] = dshp / shp * vec3[k];
            double hs = hscale(m, m.z[jat], jsh, m.z[iat], ish) * shp;
            double hij = 0.5 * (se[si] + se[sj]) * hs;
            double dhdcni = dsedcn[si] * hs, dhdcnj = dsedcn[sj] * hs;
            int nj = nsph(cj.l);
            for (int iao = 0; iao < nsph(ci.l); iao++)
              for (int jao = 0; jao < nj; jao++) {
                int ij = jao + nj * iao;
                int mu = m.iao_sh[sj] + jao, nu = m.iao_sh[si] + iao;
                double pij = P(mu, nu);
                double hpij = pij * hij;
                double sval = -pij * (vao[mu] + vao[nu]) + 2.0 * hpij - 2.0 * W(mu, nu);
                for (int k = 0; k < 3; k++) {
                  double mp = 0.0;
                  for (int c = 0; c < 3; c++)
                    mp += mg.ddi[ij][k][c] * pot.vdp[3 * iat + c] + mg.ddj[ij][k][c] * pot.vdp[3 * jat + c];
                  for (int c = 0; c < 6; c++)
                    mp += mg.dqi[ij][k][c] * pot.vqp[6 * iat + c] + mg.dqj[ij][k][c] * pot.vqp[6 * jat + c];
                  dg[k] += sval * mg.ds[ij][k] + 2.0 * hpij * mg.s[ij] * dsv[k] - pij * mp;
                }
                dcni += dhdcni * pij * mg.s[ij];
                dcnj += dhdcnj * pij * mg.s[ij];
              }
          }
        }
        dedcn[iat] += dcni;
        dedcn[jat] += dcnj;
        for (int k = 0; k < 3; k++) {
          g[3 * iat + k] += dg[k];
          g[3 * jat + k] -= dg[k];
        }
      }
      for (int ish = 0; ish < m.nsh_at[iat]; ish++) {
        int s = m.ish_at[iat] + ish;
        int ii = m.iao_sh[s];
        double d = 0.0;
        for (int iao = 0; iao < m.nao_sh[s]; iao++) d += dsedcn[s] * P(ii + iao, ii + iao);
        dedcn[iat] += d;
      }
    }
    }
    for (int t = 0; t < nt; t++) {
      for (int a = 0; a < 3 * nat; a++) g[a] += tg[t][a];
      for (int a = 0; a < nat; a++) dedcn[a] += tc[t][a];
    }
    // CN chain (terms/chain.rs:288, base/src/ncoord.rs:486-519)
    gfn2_cn_gradient(m, dedcn, g);
    res.grad = g;
  }
  return res;
}

thread_local std::string last_error;

}  // namespace

extern "C" {

// Sizes of the basis of a molecule: nsh, nao. Returns 0 or an error code.
int oracle_sizes(int nat, const int* z, int* nsh, int* nao) {
  try {
    double xyz0[3] = {0, 0, 0};
    (void)xyz0;
    int s = 0, a = 0;
    for (int i = 0; i < nat; i++) {
      if (z[i] < 1 || z[i] > AXTB_GFN2_MAX_ELEM) return 4;
      for (int k = 0; k < GFN2_NSHELL[z[i] - 1]; k++) {
        s++;
        a += 2 * GFN2_ANG_SHELL[z[i] - 1][k] + 1;
      }
    }
    *nsh = s;
    *nao = a;
    return 0;
  } catch (...) {
    return 1;
  }
}

const char* oracle_last_error() { return last_error.c_str(); }

// GFN2-xTB single point. Outputs (caller-owned, may be null): energy, terms[8], grad[3N], qat[N],
// qsh[nsh], dp[3N], qp[6N], emo[nao], iterations. Returns 0 ok, 1 input, 2 numerical,
// 3 SCF not converged, 4 unsupported element.
int oracle_engrad_guess(int nat, const int* z, const double* xyz, double charge, double acc, double etemp,
                        int max_iter, int guess, int want_grad, int dispersion, const double* gq, const double* gdp,
                        const double* gqp, double* energy, double* terms, double* grad, double* qat, double* qsh,
                        double* dp, double* qp, double* emo, int* iterations) {
  try {
    Options o;
    o.gq = gq;
    o.gdp = gdp;
    o.gqp = gqp;
    o.acc = acc;
    o.etemp = etemp;
    o.max_iter = max_iter;
    o.guess = guess;
    o.want_grad = want_grad;
    o.dispersion = dispersion;
    Result r = singlepoint(nat, z, xyz, charge, o);
    if (energy) *energy = r.energy;
    if (terms) std::memcpy(terms, r.terms, sizeof r.terms);
    if (grad && want_grad) std::memcpy(grad, r.grad.data(), sizeof(double) * 3 * nat);
    if (qat) std::memcpy(qat, r.qat.data(), sizeof(double) * nat);
    if (qsh) std::memcpy(qsh, r.qsh.data(), sizeof(double) * r.nsh);
    if (dp) std::memcpy(dp, r.dp.data(), sizeof(double) * 3 * nat);
    if (qp) std::memcpy(qp, r.qp.data(), sizeof(double) * 6 * nat);
    if (emo) std::memcpy(emo, r.emo.data(), sizeof(double) * r.nao);
    if (iterations) *iterations = r.iterations;
    return 0;
  } catch (Error& e) {
    last_error = e.what();
    return e.code;
  } catch (std::exception& e) {
    last_error = e.what();
    return 2;
  }
}

// The single point with the mixer plan named (0 broyden, 1 diis, 2 auto = the axtb conventions'
// default) and the SCF report: info[0..5] = iterations of the result (of all attempts on
// ScfNotConverged), mixer used (0 broyden, 1 diis, 2 diis-fallback), attempts, attempt 1
// iterations, attempt 1 converged, attempt 2 iterations; info[6] = attempt 2 converged.
// scal[0..4] = efermi, dipole x y z, HOMO-LUMO gap.
// Implicit solvation: solv_model 0 none, 1 ALPB, 2 GBSA; the solvent's name; solv_state 0 gsolv,
// 1 bar1M, 2 reference. terms[9]: ... band, solvation.
int oracle_singlepoint(int nat, const int* z, const double* xyz, double charge, double acc, double etemp,
                       int max_iter, int guess, int want_grad, int dispersion, int mixer, const double* gq,
                       const double* gdp, const double* gqp, double* energy, double* terms, double* grad, double* qat,
                       double* qsh, double* dp, double* qp, double* emo, int* info, double* scal, int solv_model,
                       const char* solvent, int solv_state, int conventions, int uhf) {
  Options o;
  o.conventions = conventions;
  o.uhf = uhf;
  o.solv_model = solv_model;
  o.solvent = solvent;
  o.solv_state = solv_state;
  o.gq = gq;
  o.gdp = gdp;
  o.gqp = gqp;
  o.acc = acc;
  o.etemp = etemp;
  o.max_iter = max_iter;
  o.guess = guess;
  o.want_grad = want_grad;
  o.dispersion = dispersion;
  o.mixer = mixer;
  Result r;
  try {
    r = singlepoint(nat, z, xyz, charge, o, &r);
  } catch (Error& e) {
    last_error = e.what();
    if (info) {
      info[0] = r.iterations;
      info[1] = r.mixer_used;
      info[2] = r.nattempts;
      info[3] = r.attempt_iters[0];
      info[4] = r.attempt_conv[0];
      info[5] = r.attempt_iters[1];
      info[6] = r.attempt_conv[1];
    }
    return e.code;
  } catch (std::exception& e) {
    last_error = e.what();
    return 2;
  }
  if (energy) *energy = r.energy;
  if (terms) std::memcpy(terms, r.terms, sizeof r.terms);
  if (grad && want_grad) std::memcpy(grad, r.grad.data(), sizeof(double) * 3 * nat);
  if (qat) std::memcpy(qat, r.qat.data(), sizeof(double) * nat);
  if (qsh) std::memcpy(qsh, r.qsh.data(), sizeof(double) * r.nsh);
  if (dp) std::memcpy(dp, r.dp.data(), sizeof(double) * 3 * nat);
  if (qp) std::memcpy(qp, r.qp.data(), sizeof(double) * 6 * nat);
  if (emo) std::memcpy(emo, r.emo.data(), sizeof(double) * r.nao);
  if (info) {
    info[0] = r.iterations;
    info[1] = r.mixer_used;
    info[2] = r.nattempts;
    info[3] = r.attempt_iters[0];
    info[4] = r.attempt_conv[0];
    info[5] = r.attempt_iters[1];
    info[6] = r.attempt_conv[1];
  }
  if (scal) {
    scal[0] = r.efermi;
    for (int k = 0; k < 3; k++) scal[1 + k] = r.dipole[k];
    scal[4] = r.gap;
  }
  return 0;
}

// The solvation model alone at given charges (positions in bohr): out[5] = gsolv, gelec, gsasa, ghb,
// gshift as xtb prints them; Born radii and surface areas (bohr^2 / 4 pi) per atom if not null; the
// gradient of gsolv at fixed charges (3N) if not null. conv: 0 xtb 6.8.0, 1 xtb 6.7.1, 2 corrected.
int oracle_solv_parts(int nat, const int* z, const double* xyz, const double* q, int model, const char* solvent,
                      int state, int conv, int nang, double* out, double* brad, double* sasa, double* grad) {
  try {
    axtb_solv::SolvModel S;
    std::string err;
    if (!axtb_solv::build_model(nat, z, model, solvent, state, conv, nang, axtb_solv::DEFAULT_TEMPERATURE, S, err))
      throw Error(1, err);
    vec x(xyz, xyz + 3 * nat), qq(q, q + nat);
    SolvCache c;
    solv_update(S, x, c);
    solv_parts(S, c, qq, out);
    if (brad) std::memcpy(brad, c.brad.data(), sizeof(double) * nat);
    if (sasa) std::memcpy(sasa, c.sasa.data(), sizeof(double) * nat);
    if (grad) {
      vec g(3 * nat, 0.0);
      solv_gradient(S, x, c, qq, g);
      solv_sasa_contract(c, S.gamsasa, g);
      std::memcpy(grad, g.data(), sizeof(double) * 3 * nat);
    }
    return 0;
  } catch (Error& e) {
    last_error = e.what();
    return e.code;
  }
}

// Many molecules at once, one molecule per OpenMP thread (inner loops serial): the CPU baseline
// of a batch, as axtb runs many calculations (one thread each). Molecule i has nats[i] atoms at
// offset off[i] of z / xyz (bohr); energies[i] and status[i] are written; gradients (3 N_i at
// 3 off[i]) if grads is not null.
int oracle_engrad_many(int nmol, const int* nats, const long long* off, const int* z, const double* xyz,
                       double acc, double etemp, int want_grad, double* energies, double* grads, int* status) {
  const int outer = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(outer)
  for (int i = 0; i < nmol; i++) {
    omp_set_num_threads(1);
    Options o;
    o.acc = acc;
    o.etemp = etemp;
    o.want_grad = want_grad;
    try {
      Result r = singlepoint(nats[i], z + off[i], xyz + 3 * off[i], 0.0, o);
      energies[i] = r.energy;
      if (grads && want_grad) std::memcpy(grads + 3 * off[i], r.grad.data(), sizeof(double) * 3 * nats[i]);
      status[i] = 0;
    } catch (Error& e) {
      status[i] = e.code;
    } catch (std::exception&) {
      status[i] = 2;
    }
  }
  return 0;
}

// The DIIS mixer alone on the linear fixed point q = diag(eig) q + b (the reference's own test,
// scf/mixing.rs:523-546): iterates until the RMS residual is below 1e-12 or max_steps steps;
// returns the steps taken (-1: the mixer failed), the final input in x.
int oracle_diis_linear(int n, const double* eig, const double* b, double damp, int max_steps, double* x) {
  Diis m(n, damp);
  vec out(n);
  for (int it = 0; it <= max_steps; it++) {
    for (int k = 0; k < n; k++) out[k] = eig[k] * m.q_in[k] + b[k];
    for (int k = 0; k < n; k++) m.dq[k] = out[k] - m.q_in[k];
    if (std::sqrt(dot_seq(m.dq, m.dq) / n) < 1e-12) {
      std::memcpy(x, m.q_in.data(), sizeof(double) * n);
      return it;
    }
    if (it == max_steps) break;
    if (!m.next()) return -1;
  }
  std::memcpy(x, m.q_in.data(), sizeof(double) * n);
  return max_steps + 1;
}

// Pulay coefficients of m residuals of length n (row i = residual i, oldest first) into c[m]
// (scf/mixing.rs:467-513).
void oracle_diis_coefficients(int m, int n, const double* r, double* c) {
  std::deque<vec> h;
  for (int i = 0; i < m; i++) h.push_back(vec(r + (size_t)i * n, r + (size_t)(i + 1) * n));
  vec cc = diis_coefficients(h);
  std::memcpy(c, cc.data(), sizeof(double) * m);
}

// OpenMP threads of the following calls (the batch entry: molecules at once; one molecule: its loops).
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

int oracle_engrad(int nat, const int* z, const double* xyz, double charge, double acc, double etemp, int max_iter,
                  int guess, int want_grad, int dispersion, double* energy, double* terms, double* grad, double* qat,
                  double* qsh, double* dp, double* qp, double* emo, int* iterations) {
  return oracle_engrad_guess(nat, z, xyz, charge, acc, etemp, max_iter, guess, want_grad, dispersion, nullptr, nullptr,
                             nullptr, energy, terms, grad, qat, qsh, dp, qp, emo, iterations);
}

// EEQ-2019 charges of the SCF start (axtb/src/eeq.rs:43-77)
int oracle_eeq(int nat, const int* z, const double* xyz, double charge, double* q) {
  try {
    vec r = eeq_charges(nat, z, xyz, charge);
    std::memcpy(q, r.data(), sizeof(double) * nat);
    return 0;
  } catch (Error& e) {
    last_error = e.what();
    return e.code;
  }
}

// D4 pieces for the fixtures: cn[N], c6[N*N], alpha[N], e2, e3 (sums), grad[3N] (2-body at fixed q
// + 3-body, incl. the D4 CN chain). cutoffs = {cn, disp2, disp3, width2, width3}.
int oracle_d4(int nat, const int* z, const double* xyz, const double* q, int charge_model, int rcov_mode,
              const double* cutoffs, double* cn, double* c6, double* alpha, double* e2, double* e3, double* grad) {
  try {
    D4Options o;
    o.charge_model = charge_model;
    o.rcov_mode = rcov_mode;
    o.cn_cut = cutoffs[0];
    o.disp2 = cutoffs[1];
    o.disp3 = cutoffs[2];
    o.width2 = cutoffs[3];
    o.width3 = cutoffs[4];
    D4 d = build_d4(nat, z, xyz, o);
    vec qq(q, q + nat);
    if (cn) std::memcpy(cn, d.cn.data(), sizeof(double) * nat);
    vector<vec> gw, dgwdcn, dgwdq;
    d4_weights(d, qq, gw, &dgwdcn, &dgwdq);
    if (c6)
      for (int i = 0; i < nat; i++)
        for (int j = 0; j < nat; j++) {
          double c, a, b;
          d4_pair_c6(d, i, j, gw[i], dgwdq[i], gw[j], dgwdq[j], c, a, b);
          c6[i * nat + j] = c;
        }
    if (alpha)
      for (int i = 0; i < nat; i++) {
        alpha[i] = 0.0;
        for (int r = 0; r < d.nref[i]; r++) alpha[i] += gw[i][r] * d.aiw[i][r][0];
      }
    vec ea(nat, 0.0), eb(nat, 0.0);
    d4_energy_potential(d, qq, &ea, nullptr);
    vec g(3 * nat, 0.0);
    d4_three_body(d, eb, grad ? &g : nullptr);
    if (grad) d4_gradient(d, qq, g);
    double s2 = 0, s3 = 0;
    for (int i = 0; i < nat; i++) {
      s2 += ea[i];
      s3 += eb[i];
    }
    if (e2) *e2 = s2;
    if (e3) *e3 = s3;
    if (grad) std::memcpy(grad, g.data(), sizeof(double) * 3 * nat);
    return 0;
  } catch (Error& e) {
    last_error = e.what();
    return e.code;
  }
}

// Overlap matrix (column-major nao x nao) and the GFN2 CN of a molecule, for integral checks.
int oracle_overlap(int nat, const int* z, const double* xyz, double acc, double* S, double* cn_out) {
  try {
    Model m = build_model(nat, z, xyz, 0.0, acc);
    vec cn = gfn2_cn(m);
    vec se(m.nsh, 0.0);
    Ints I = build_integrals(m, se);
  {
    size_t nn = (size_t)m.nao * m.nao;
    dbg_copy(debug_out.S, I.S.a.data(), nn);
    dbg_copy(debug_out.H0, I.H0.a.data(), nn);
    for (int k = 0; k < 3; k++) dbg_copy(debug_out.D ? debug_out.D + k * nn : nullptr, I.D[k].a.data(), nn);
    for (int k = 0; k < 6; k++) dbg_copy(debug_out.Q ? debug_out.Q + k * nn : nullptr, I.Q[k].a.data(), nn);
    dbg_copy(debug_out.cn, cn.data(), nat);
  }
    if (S) std::memcpy(S, I.S.a.data(), sizeof(double) * m.nao * m.nao);
    if (cn_out) std::memcpy(cn_out, cn.data(), sizeof(double) * nat);
    return 0;
  } catch (Error& e) {
    last_error = e.what();
    return e.code;
  }
}

// Dense symmetric eigensolver of the oracle (for its own residual checks): a overwritten by
// the eigenvectors (column-major), w ascending.
int oracle_eigh(int n, double* a, double* w) {
  try {
    vec A(a, a + (size_t)n * n), W;
    eigh(n, A, W);
    std::memcpy(a, A.data(), sizeof(double) * n * n);
    std::memcpy(w, W.data(), sizeof(double) * n);
    return 0;
  } catch (Error& e) {
    last_error = e.what();
    return e.code;
  }
}


// Setup and first SCC iteration of the oracle for kernel-level parity checks: S, H0 (nao^2
// column-major), D (3 nao^2), Q (6 nao^2), GFN2 CN, start shell charges, then the first iteration's
// potentials (vsh, vat, vdp, vqp), Fock matrix F, eigenvalues w, density P and output shell charges.
int oracle_debug(int nat, const int* z, const double* xyz, double charge, double acc, double* S, double* H0,
                 double* D, double* Q, double* cn_out, double* qstart, double* vsh, double* vat, double* vdp,
                 double* vqp, double* F, double* w, double* P, double* qsh_out) {
  try {
    Options opt;
    opt.acc = acc;
    opt.max_iter = 1;
    debug_out = DebugOut{S, H0, D, Q, cn_out, qstart, vsh, vat, vdp, vqp, F, w, P, qsh_out};
    debug_on = true;
    try {
      singlepoint(nat, z, xyz, charge, opt);
    } catch (Error& e) {
      if (e.code != 3) throw;
    }
    debug_on = false;
    return 0;
  } catch (Error& e) {
    debug_on = false;
    last_error = e.what();
    return e.code;
  }
}

}  // extern "C"
