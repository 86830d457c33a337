// ALPB / GBSA implicit solvation, host side: the model's constants, the solvent lookup, the
// Lebedev-Laikov angular grids and the per-atom parameters of one molecule. Shared by the engine
// (which uploads them for its kernels) and the CPU oracle; the Born radii, surface areas, Born
// matrix, energies and gradients are computed by each on its own.
// Reference: crates/solvation/src/lib.rs (model setup 380-497, reference-state shift 282-312),
// params.rs:9-51 (constants), lebedev.rs (grids).
#pragma once

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "solv_tables.h"

namespace axtb_solv {

// xtb's constants of the model (crates/solvation/src/params.rs:9-51)
constexpr int DEFAULT_NANG = 230;
constexpr double XTB_AUTOAA = 0.52917726;
constexpr double AUTOKCAL = 627.50947428;
constexpr double PI = 3.14159265358979323846;
constexpr double KB_XTB = 3.166808578545117e-06;
constexpr double MOLCM3TOAU = 8.92388e-2;
constexpr double SURFACE_TENSION = 1.0e-5;
constexpr double LRCUT_AA = 35.0;
constexpr double SRCUT_AA = 2.0;
constexpr double REF_DENSITY = 1.0e-3;
constexpr double REF_MOLECULAR_MASS = 1.0;
constexpr double IDEAL_GAS_MOL_VOLUME_XTB = 24.79;
constexpr double AMBIENT_TEMPERATURE = 298.15;
constexpr double DEFAULT_TEMPERATURE = 298.15;
constexpr double ALPB_ALPHA = 0.571412;
constexpr double SMOOTH_W_AA = 0.3;
constexpr double TOLSESP = 1.0e-6;
constexpr double ZETA_P16 = 1.028;
constexpr double OBC_ALPHA = 1.0;
constexpr double OBC_BETA = 0.8;
constexpr double OBC_GAMMA = 4.85;
// CODATA 2022 values of the corrected conventions (crates/base/src/units.rs:18-37, solvation
// lib.rs:128): bohr, kB, Hartree in kcal/mol, and the ideal-gas molar volume at 298.15 K, 1 bar
// (R T / p, in litres, in the reference's operation order)
constexpr double AUTOAA_2022 = 5.29177210544e-11 * 1e10;
constexpr double KB_HARTREE_2022 = 3.1668115634564e-6;
constexpr double HARTREE_TO_KCAL_2022 = 4.3597447222060e-18 * 6.02214076e23 / 4184.0;
constexpr double IDEAL_GAS_MOL_VOLUME_2022 = 6.02214076e23 * 1.380649e-23 * 298.15 / 1.0e5 * 1.0e3;
// the Angstrom-to-bohr factors of xtb 6.7.1 and 6.8.0 (units.rs:84-88)
constexpr double XTB671_AATOAU = 1.0 / 0.52917726;
constexpr double XTB680_AATOAU = 1.0 / 5.29177210904492434e-1;

enum Model { NONE = 0, ALPB = 1, GBSA = 2 };
enum State { GSOLV = 0, BAR1M = 1, REFERENCE = 2 };
// numerical conventions (lib.rs:100-149): xtb 6.8.0, xtb 6.7.1 (its Born-radius quirk and its
// reference-state density), corrected (CODATA 2022; the axtb conventions, conventions.rs:256)
enum Conv { CONV_XTB = 0, CONV_XTB671 = 1, CONV_CORRECTED = 2 };

// parameter set of `solvent` for `model`: xtb's lower-casing, aliases and GFN2 selection table
// (lib.rs:255-271); -1 when the model has no parameters for it
inline int find_set(const char* solvent, int model) {
  if (!solvent || (model != ALPB && model != GBSA)) return -1;
  std::string s(solvent);
  const auto b = s.find_first_not_of(" \t\r\n"), e = s.find_last_not_of(" \t\r\n");
  s = b == std::string::npos ? std::string() : s.substr(b, e - b + 1);
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  std::string norm = s;
  for (int k = 0; k < SOLV_NALIAS; k++)
    if (s == SOLV_ALIASES[k].alias) {
      norm = SOLV_ALIASES[k].target;
      break;
    }
  const SolvName* tab = model == ALPB ? SOLV_ALPB : SOLV_GBSA;
  const int nt = model == ALPB ? SOLV_ALPB_N : SOLV_GBSA_N;
  for (int k = 0; k < nt; k++)
    if (norm == tab[k].name) return tab[k].set;
  return -1;
}

// xtb's bisection on the grid sizes (lebedev.rs:32-56): the 1-based grid j with
// size[j] <= n < size[j + 1], the largest size mapping to the second largest grid
inline int bisect_grid(int n) {
  const int len = LEB_NGRID;
  const int* xx = LEB_GRID_SIZE;
  if (n < xx[0]) return 0;
  if (n == xx[0]) return 1;
  if (n == xx[len - 1]) return len - 1;
  if (n > xx[len - 1]) return len;
  int jlower = 0, jcurr = len + 1;
  while (jcurr - jlower > 1) {
    const int jupper = (jcurr + jlower) / 2;
    if (n >= xx[jupper - 1])
      jlower = jupper;
    else
      jcurr = jupper;
  }
  return jlower;
}

// Points (x, y, z) and weights of grid iang (1-based), in the Lebedev-Laikov order of each Oh
// generator: the coordinate permutations in turn, each with its sign patterns (the first nonzero
// coordinate's sign alternating fastest). False for an invalid index.
inline bool angular_grid(int iang, std::vector<double>& xyz, std::vector<double>& w) {
  xyz.clear();
  w.clear();
  if (iang < 1 || iang > LEB_NGRID) return false;
  auto put = [&](double x, double y, double z, double v) {
    xyz.push_back(x);
    xyz.push_back(y);
    xyz.push_back(z);
    w.push_back(v);
  };
  // all sign patterns of the nonzero entries of p, the first nonzero entry's sign fastest
  auto signs = [&](const double p[3], double v) {
    int nz[3], k = 0;
    for (int c = 0; c < 3; c++)
      if (p[c] != 0.0) nz[k++] = c;
    for (int s = 0; s < (1 << k); s++) {
      double q[3] = {p[0], p[1], p[2]};
      for (int t = 0; t < k; t++)
        if (s >> t & 1) q[nz[t]] = -q[nz[t]];
      put(q[0], q[1], q[2], v);
    }
  };
  for (int g = LEB_GEN_OFF[iang - 1]; g < LEB_GEN_OFF[iang]; g++) {
    const LebGen& G = LEB_GEN[g];
    const double v = G.v;
    switch (G.kind) {
      case 1: {
        const double a = 1.0;
        put(a, 0, 0, v);
        put(-a, 0, 0, v);
        put(0, a, 0, v);
        put(0, -a, 0, v);
        put(0, 0, a, v);
        put(0, 0, -a, v);
        break;
      }
      case 2: {
        const double a = std::sqrt(0.5);
        const double p[3][3] = {{0, a, a}, {a, 0, a}, {a, a, 0}};
        for (auto& q : p) signs(q, v);
        break;
      }
      case 3: {
        const double a = std::sqrt(1.0 / 3.0);
        const double p[3] = {a, a, a};
        signs(p, v);
        break;
      }
      case 4: {
        const double a = G.a, b = std::sqrt(1.0 - 2.0 * a * a);
        const double p[3][3] = {{a, a, b}, {a, b, a}, {b, a, a}};
        for (auto& q : p) signs(q, v);
        break;
      }
      case 5: {
        const double a = G.a, b = std::sqrt(1.0 - a * a);
        const double p[6][3] = {{a, b, 0}, {b, a, 0}, {a, 0, b}, {b, 0, a}, {0, a, b}, {0, b, a}};
        for (auto& q : p) signs(q, v);
        break;
      }
      case 6: {
        const double a = G.a, b = G.b, c = std::sqrt(1.0 - a * a - b * b);
        const double p[6][3] = {{a, b, c}, {a, c, b}, {b, a, c}, {b, c, a}, {c, a, b}, {c, b, a}};
        for (auto& q : p) signs(q, v);
        break;
      }
      default:
        return false;
    }
  }
  return (int)w.size() == LEB_GRID_SIZE[iang - 1];
}

// reference-state shift in Hartree with a given kB (Eh/K) and ideal-gas molar volume (L)
// (lib.rs:290-312)
inline double state_shift_with(int state, double temperature, double density, double molar_mass, double kb,
                               double vm) {
  if (state == REFERENCE)
    return temperature * kb *
           (std::log(vm * temperature / AMBIENT_TEMPERATURE) +
            std::log(density / REF_DENSITY * REF_MOLECULAR_MASS / molar_mass));
  if (state == BAR1M) return temperature * kb * std::log(vm * temperature / AMBIENT_TEMPERATURE);
  return 0.0;
}

// total free-energy shift (empirical + reference state) of a parameter set (lib.rs:150-170)
inline double free_energy_shift(int conv, const SolvParamSet& p, int state, double temperature) {
  const double kcaltoau = 1.0 / AUTOKCAL;
  if (conv == CONV_XTB671) {
    const double density = p.rhos / p.smass * MOLCM3TOAU;
    return p.gshift * kcaltoau +
           state_shift_with(state, temperature, density, p.smass, KB_XTB, IDEAL_GAS_MOL_VOLUME_XTB);
  }
  if (conv == CONV_XTB)
    return p.gshift * kcaltoau + state_shift_with(state, temperature, p.rhos, p.smass, KB_XTB, IDEAL_GAS_MOL_VOLUME_XTB);
  return p.gshift / HARTREE_TO_KCAL_2022 +
         state_shift_with(state, temperature, p.rhos, p.smass, KB_HARTREE_2022, IDEAL_GAS_MOL_VOLUME_2022);
}

// The model for one molecule (lib.rs:380-497): scalars and per-atom parameters, bohr and Hartree.
struct SolvModel {
  int model = NONE, conv = CONV_CORRECTED, set = -1, state = GSOLV;
  bool alpb = false, fresh_r1 = true, lhb = false;
  double aatoau = 0, dielectric = 0, keps = 0, alpbet = 0, born_scale = 0, born_offset = 0, probe = 0;
  double gshift = 0, lrcut = 0, srcut = 0, w = 0, ah0 = 0, ah1 = 0, ah3 = 0;
  // per atom: vdW radius, descreened radius rho, vdW radius minus the Born offset, SASA radius,
  // squared inner and outer smoothing radii, the surface normalisation wrp, surface tension
  // (4 pi gamma), H-bond strength (Eh)
  std::vector<double> vdwr, rho, svdw, vdwsa, trj2a, trj2b, wrp, gamsasa, hbmag;
  std::vector<double> grid, weight;  // angular grid: points (x, y, z) and weights
};

// Build the model of a molecule (numbers 1..94). Returns false with a message for an unknown
// solvent, an element without parameters or an invalid grid.
inline bool build_model(int nat, const int* z, int model, const char* solvent, int state, int conv, int nang,
                        double temperature, SolvModel& S, std::string& err) {
  S = SolvModel();
  if (model != ALPB && model != GBSA) {
    err = "unknown solvation model";
    return false;
  }
  if (state < GSOLV || state > REFERENCE) {
    err = "unknown solvation reference state";
    return false;
  }
  for (int i = 0; i < nat; i++)
    if (z[i] < 1 || z[i] > SOLV_NELEM) {
      err = "solvation: no parameters for atomic number " + std::to_string(z[i]);
      return false;
    }
  S.set = find_set(solvent, model);
  if (S.set < 0) {
    err = std::string("solvent '") + (solvent ? solvent : "") + "' is not parametrized for GFN2-xTB/" +
          (model == ALPB ? "ALPB" : "GBSA");
    return false;
  }
  const SolvParamSet& p = SOLV_SETS[S.set];
  S.model = model;
  S.conv = conv;
  S.state = state;
  S.alpb = model == ALPB;
  S.fresh_r1 = conv != CONV_XTB671;
  S.aatoau = conv == CONV_XTB671 ? XTB671_AATOAU : conv == CONV_XTB ? XTB680_AATOAU : 1.0 / AUTOAA_2022;
  const double aatoau = S.aatoau;
  S.dielectric = p.epsv;
  S.born_scale = p.c1;
  S.born_offset = p.soset * 0.1 * aatoau;
  S.probe = p.rprobe * aatoau;
  bool any_hb = false;
  for (int k = 0; k < SOLV_NELEM; k++) any_hb |= std::fabs(p.tmp[k]) > 1.0e-3;
  S.gshift = free_energy_shift(conv, p, state, temperature);
  S.alpbet = S.alpb ? ALPB_ALPHA / S.dielectric : 0.0;
  S.keps = (1.0 / S.dielectric - 1.0) / (1.0 + S.alpbet);
  S.lrcut = LRCUT_AA * aatoau;
  const double w = SMOOTH_W_AA * aatoau, w3 = w * (w * w);
  S.w = w;
  S.ah0 = 0.5;
  S.ah1 = 3.0 / (4.0 * w);
  S.ah3 = -1.0 / (4.0 * w3);
  const double fourpi = 4.0 * PI, kcaltoau = 1.0 / AUTOKCAL;
  double maxsa = -INFINITY;
  for (int i = 0; i < nat; i++) {
    const int iz = z[i] - 1;
    const double rvdw = aatoau * SOLV_VDW_RAD_D3_AA[iz];
    S.vdwr.push_back(rvdw);
    S.rho.push_back(rvdw * p.sx[iz]);
    S.svdw.push_back(rvdw - S.born_offset);
    const double rsa = rvdw + S.probe;
    S.vdwsa.push_back(rsa);
    S.trj2a.push_back((rsa - w) * (rsa - w));
    S.trj2b.push_back((rsa + w) * (rsa + w));
    double r = rsa + w;
    double wr = (0.25 / w + 3.0 * S.ah3 * (0.2 * r * r - 0.5 * r * rsa + rsa * rsa / 3.0)) * r * r * r;
    r = rsa - w;
    wr -= (0.25 / w + 3.0 * S.ah3 * (0.2 * r * r - 0.5 * r * rsa + rsa * rsa / 3.0)) * r * r * r;
    S.wrp.push_back(wr);
    S.gamsasa.push_back(p.gamscale[iz] * fourpi * SURFACE_TENSION);
    S.hbmag.push_back(any_hb ? -(p.tmp[iz] * p.tmp[iz]) * kcaltoau : 0.0);
    maxsa = std::max(maxsa, rsa);
  }
  S.srcut = 2.0 * (w + maxsa) + SRCUT_AA * aatoau;
  for (double h : S.hbmag) S.lhb |= h < 0.0;
  if (!angular_grid(bisect_grid(nang), S.grid, S.weight)) {
    err = "no Lebedev grid for " + std::to_string(nang) + " points";
    return false;
  }
  return true;
}

}  // namespace axtb_solv
