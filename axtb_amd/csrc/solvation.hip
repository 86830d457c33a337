// ALPB / GBSA implicit solvation on the device (the axtb conventions' corrected model,
// crates/solvation/src/lib.rs; xtb's loops restated for a batch of molecules): Born radii
// (born.rs), the smoothed solvent-accessible surface (sasa.rs), the Born matrix with the P16 or
// Still kernel (kernel.rs), the per-iteration potential B q and energy q (B q) / 2, and the
// gradient at fixed charges. One wavefront per atom over its partners (sums in a fixed order);
// no atomics, so results repeat bit for bit.
#pragma once
#include <cfloat>

#include "common.hpp"
#include "solvation.hpp"

namespace axtb {

constexpr int SOLV_MAXNB = 512;  // SASA neighbours of one atom (LDS list); more: the molecule fails

struct SolvDev {
  // per atom: vdW radius, descreened radius, vdW radius minus the Born offset, SASA radius, inner
  // and outer smoothing radii squared, surface normalisation, surface tension, H-bond strength
  const double *vdwr, *rho, *svdw, *vdwsa, *trj2a, *trj2b, *wrp, *gam, *hbmag;
  const double* grid;  // angular grid: x, y, z, weight per point
  int npts, alpb;
  double keps, alpbet, c1, lrcut2, w, ah0, ah1, ah3, gshift;
  // geometry-dependent: Born radii and d brad / d psi, surface areas, point weights (npts per
  // atom), fixed energies per atom, aDet per molecule, Born matrices (row-major at pair0), and
  // dE/dbrad * dbrad/dpsi of the gradient
  double *brad, *dbrad, *sasa, *wsa, *efix, *adet, *B, *work;
};

struct BornPsi {
  double g, dg;
};

// the GBOBC sphere integrals (born.rs:20-52), r1 = 1 / r (a fresh r1 for every pair)
__device__ __forceinline__ BornPsi solv_nonoverlap(double r, double rho, double r1) {
  const double ap = r + rho, am = r - rho, ab = ap * am, rhab = rho / ab;
  const double lnab = 0.5 * log(am / ap) * r1;
  return BornPsi{rhab + lnab, -2.0 * rhab / ab + (rhab - lnab) * r1 * r1};
}
__device__ __forceinline__ BornPsi solv_overlap(double r, double rho, double rvdw, double r1) {
  const double r12 = 0.5 * r1, r24 = r12 * r12, ap = r + rho, am = r - rho;
  const double rh1 = 1.0 / rvdw, rhr1 = 1.0 / ap, aprh1 = ap * rh1, lnab = log(aprh1);
  const double g = rh1 - rhr1 + r12 * (0.5 * am * (rhr1 - rh1 * aprh1) - lnab);
  const double dg = rhr1 * rhr1 * (1.0 - 0.25 * am * r1 * (1.0 + aprh1 * aprh1)) + rho * r24 * (rhr1 - rh1 * aprh1) +
                    r12 * (r1 * lnab - rhr1);
  return BornPsi{g, dg * r1};
}
// contributions of the pair (ii > jj) to psi(ii) and psi(jj) in the four overlap cases
// (born.rs:54-100); has_i / has_j false where the case adds nothing
__device__ __forceinline__ void solv_born_pair(double r, double rhoi, double rhoj, double rvdwi, double rvdwj,
                                               BornPsi& ci, bool& hi, BornPsi& cj, bool& hj) {
  const double r1 = 1.0 / r;
  const bool ovij = !(r >= rvdwi + rhoj), ovji = !(r >= rhoi + rvdwj);
  hi = hj = false;
  ci = cj = BornPsi{0.0, 0.0};
  if (!ovij && !ovji) {
    ci = solv_nonoverlap(r, rhoj, r1);
    cj = fabs(rhoi - rhoj) < 1.0e-8 ? ci : solv_nonoverlap(r, rhoi, r1);
    hi = hj = true;
  } else if (!ovij) {
    ci = solv_nonoverlap(r, rhoj, r1);
    hi = true;
    if (r + rhoi > rvdwj) {
      cj = solv_overlap(r, rhoi, rvdwj, r1);
      hj = true;
    }
  } else if (!ovji) {
    if (r + rhoj > rvdwi) {
      ci = solv_overlap(r, rhoj, rvdwi, r1);
      hi = true;
    }
    cj = solv_nonoverlap(r, rhoi, r1);
    hj = true;
  } else {
    if (r + rhoj > rvdwi) {
      ci = solv_overlap(r, rhoj, rvdwi, r1);
      hi = true;
    }
    if (r + rhoi > rvdwj) {
      cj = solv_overlap(r, rhoi, rvdwj, r1);
      hj = true;
    }
  }
}
// the contribution of partner j to psi(a) ("mine") and of a to psi(j) ("other")
__device__ __forceinline__ void solv_born_contrib(const SolvDev& S, int a, int j, double r, BornPsi& mine, bool& hm,
                                                  BornPsi& other, bool& ho) {
  const int ii = max(a, j), jj = min(a, j);
  BornPsi ci, cj;
  bool hi, hj;
  solv_born_pair(r, S.rho[ii], S.rho[jj], S.vdwr[ii], S.vdwr[jj], ci, hi, cj, hj);
  if (a == ii) {
    mine = ci, hm = hi, other = cj, ho = hj;
  } else {
    mine = cj, hm = hj, other = ci, ho = hi;
  }
}

// Born radii (born.rs:112-160): psi of atom a summed over its partners within lrcut, then the
// OBC-II transformation. One wavefront per atom.
__global__ __launch_bounds__(256) void k_solv_born(Batch b, SolvDev S) {
  const int a = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (a >= b.nat) return;
  const Mol M = b.mol[b.a.mol[a]];
  const double* xa = &b.a.xyz[3 * a];
  double psi = 0.0;
  for (int j = M.at0 + lane; j < M.at0 + M.nat; j += 64) {
    if (j == a) continue;
    const double* xj = &b.a.xyz[3 * j];
    const double d[3] = {xa[0] - xj[0], xa[1] - xj[1], xa[2] - xj[2]};
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (!(r2 < S.lrcut2)) continue;
    BornPsi mine, other;
    bool hm, ho;
    solv_born_contrib(S, a, j, sqrt(r2), mine, hm, other, ho);
    if (hm) psi += mine.g;
  }
  psi = wave_sum(psi);
  if (lane) return;
  using namespace axtb_solv;
  double br = psi;
  const double s1 = 1.0 / S.svdw[a], v1 = 1.0 / S.vdwr[a], s2 = 0.5 * S.svdw[a];
  br *= s2;
  double arg2 = br * (OBC_GAMMA * br - OBC_BETA);
  const double arg = br * (OBC_ALPHA + arg2);
  arg2 = 2.0 * arg2 + OBC_ALPHA + OBC_GAMMA * br * br;
  const double th = tanh(arg), ch = cosh(arg);
  br = 1.0 / (s1 - v1 * th);
  br *= S.c1;
  double dpsi = ch * (s1 - v1 * th);
  dpsi = s2 * v1 * arg2 / (dpsi * dpsi);
  S.brad[a] = br;
  S.dbrad[a] = dpsi * S.c1;
}

// SASA neighbours of atom i (r < R_sa,i + R_sa,j + w, sasa.rs:34-80), ascending, into LDS: one
// wavefront of 64 lanes. Returns the count (> SOLV_MAXNB: the list is incomplete).
__device__ int solv_neighbours(const Batch& b, const SolvDev& S, int i, const Mol& M, double* nx, int* idx) {
  const int lane = threadIdx.x & 63;
  const double* xi = &b.a.xyz[3 * i];
  int cnt = 0;
  for (int j0 = M.at0; j0 < M.at0 + M.nat; j0 += 64) {
    const int j = j0 + lane;
    bool near = false;
    if (j < M.at0 + M.nat && j != i) {
      const double* xj = &b.a.xyz[3 * j];
      const double d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
      const double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      const double rmax = S.vdwsa[i] + S.vdwsa[j] + S.w + 1.0e-6;
      near = d2 < rmax * rmax;
    }
    const unsigned long long m = __ballot(near);
    const int pos = cnt + __popcll(m & ((1ull << lane) - 1ull));
    if (near && pos < SOLV_MAXNB) {
      const double* xj = &b.a.xyz[3 * j];
      nx[3 * pos] = xj[0];
      nx[3 * pos + 1] = xj[1];
      nx[3 * pos + 2] = xj[2];
      idx[pos] = j;
    }
    cnt += __popcll(m);
  }
  __syncthreads();
  return cnt;
}

// Smoothed SASA of one atom per workgroup of 64 (compute_numsa, sasa.rs:110-225): per grid point
// (lanes) the product of the neighbours' smoothing functions in ascending order, zero where any
// neighbour buries the point; the point weight wsa = w_p wr_i sasap (0 below tolsesp) is kept for
// the gradient. Also the atom's fixed energy sasa gamma + gshift / nat. An atom with more than
// SOLV_MAXNB neighbours sets fail[molecule].
__global__ __launch_bounds__(64) void k_solv_sasa(Batch b, SolvDev S, int* fail) {
  __shared__ double nx[3 * SOLV_MAXNB];
  __shared__ int idx[SOLV_MAXNB];
  const int i = blockIdx.x, lane = threadIdx.x;
  const int m = b.a.mol[i];
  const Mol M = b.mol[m];
  const int nno = solv_neighbours(b, S, i, M, nx, idx);
  double* wsa = S.wsa + (long long)i * S.npts;
  if (nno > SOLV_MAXNB) {
    for (int p = lane; p < S.npts; p += 64) wsa[p] = 0.0;
    if (lane == 0) {
      fail[m] = 1;
      S.sasa[i] = 0.0;
      S.efix[i] = 0.0;
    }
    return;
  }
  const double* xi = &b.a.xyz[3 * i];
  const double rsas = S.vdwsa[i], wr = S.wrp[i];
  double sum = 0.0;
  for (int p = lane; p < S.npts; p += 64) {
    const double* g = &S.grid[4 * p];
    const double xp[3] = {xi[0] + rsas * g[0], xi[1] + rsas * g[1], xi[2] + rsas * g[2]};
    bool buried = false;
    for (int sl = 0; sl < nno && !buried; sl++) {
      const double d[3] = {xp[0] - nx[3 * sl], xp[1] - nx[3 * sl + 1], xp[2] - nx[3 * sl + 2]};
      buried = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] <= S.trj2a[idx[sl]];
    }
    double sasap = 0.0;
    if (!buried) {
      sasap = 1.0;
      for (int sl = 0; sl < nno; sl++) {
        const int ia = idx[sl];
        const double d[3] = {xp[0] - nx[3 * sl], xp[1] - nx[3 * sl + 1], xp[2] - nx[3 * sl + 2]};
        const double tj2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (tj2 < S.trj2b[ia]) {
          const double uj = sqrt(tj2) - S.vdwsa[ia], ah3uj2 = S.ah3 * uj * uj;
          double sij = S.ah0 + (S.ah1 + ah3uj2) * uj;
          if (1.0 < sij) sij = 1.0;
          if (sij < 0.0) sij = 0.0;
          if (sij < DBL_MIN) {
            sasap = 0.0;
            break;
          }
          sasap *= sij;
        }
      }
    }
    const double ws = sasap > axtb_solv::TOLSESP ? g[3] * wr * sasap : 0.0;
    wsa[p] = ws;
    sum += ws;
  }
  sum = wave_sum(sum);
  if (lane == 0) {
    S.sasa[i] = sum;
    S.efix[i] = sum * S.gam[i] + S.gshift / M.nat;
  }
}

// ALPB shape descriptor aDet of each molecule (kernel.rs:160-209): the inertia tensor of the
// atoms weighted by their vdW radius cubed. One workgroup per molecule.
__device__ void solv_inertia(const Batch& b, const SolvDev& S, const Mol& M, double* red, double I[3][3],
                             double c[3], double& tot) {
  double t = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
  for (int a = M.at0 + threadIdx.x; a < M.at0 + M.nat; a += blockDim.x) {
    const double r = S.vdwr[a], r3 = r * r * r;
    t += r3;
    cx += b.a.xyz[3 * a] * r3;
    cy += b.a.xyz[3 * a + 1] * r3;
    cz += b.a.xyz[3 * a + 2] * r3;
  }
  tot = block_sum(t, red);
  c[0] = block_sum(cx, red) / tot;
  c[1] = block_sum(cy, red) / tot;
  c[2] = block_sum(cz, red) / tot;
  double part[6] = {0, 0, 0, 0, 0, 0};  // xx yy zz xy xz yz
  for (int a = M.at0 + threadIdx.x; a < M.at0 + M.nat; a += blockDim.x) {
    const double r = S.vdwr[a], r2 = r * r, r3 = r2 * r;
    const double v[3] = {b.a.xyz[3 * a] - c[0], b.a.xyz[3 * a + 1] - c[1], b.a.xyz[3 * a + 2] - c[2]};
    const double diag = r3 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 0.4 * r2);
    part[0] += diag - r3 * v[0] * v[0];
    part[1] += diag - r3 * v[1] * v[1];
    part[2] += diag - r3 * v[2] * v[2];
    part[3] -= r3 * v[0] * v[1];
    part[4] -= r3 * v[0] * v[2];
    part[5] -= r3 * v[1] * v[2];
  }
  double s[6];
  for (int k = 0; k < 6; k++) s[k] = block_sum(part[k], red);
  I[0][0] = s[0];
  I[1][1] = s[1];
  I[2][2] = s[2];
  I[0][1] = I[1][0] = s[3];
  I[0][2] = I[2][0] = s[4];
  I[1][2] = I[2][1] = s[5];
}
__device__ __forceinline__ double solv_det3(const double a[3][3]) {
  return a[0][0] * a[1][1] * a[2][2] - a[0][0] * a[1][2] * a[2][1] - a[0][1] * a[1][0] * a[2][2] +
         a[0][1] * a[1][2] * a[2][0] + a[0][2] * a[1][0] * a[2][1] - a[0][2] * a[1][1] * a[2][0];
}
__global__ __launch_bounds__(256) void k_solv_adet(Batch b, SolvDev S) {
  __shared__ double red[16];
  const Mol M = b.mol[blockIdx.x];
  double I[3][3], c[3], tot;
  solv_inertia(b, S, M, red, I, c, tot);
  if (threadIdx.x == 0) S.adet[blockIdx.x] = sqrt(pow(solv_det3(I), 1.0 / 3.0) / (0.4 * tot));
}

// Born matrix B (row-major nat x nat at pair0): keps / f_GB with the P16 (ALPB) or Still (GBSA)
// kernel, keps / a_i + 2 hbw_i on the diagonal (hbw_i = hbmag_i sasa_i / R_sa,i^2), plus the ALPB
// shape term keps alpbet / aDet everywhere (kernel.rs:28-70, lib.rs:598-612).
__global__ __launch_bounds__(256) void k_solv_bmat(Batch b, SolvDev S) {
  const int m = blockIdx.y;
  const Mol M = b.mol[m];
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)M.nat * M.nat) return;
  const int li = (int)(e / M.nat), lj = (int)(e % M.nat), i = M.at0 + li, j = M.at0 + lj;
  const double shift = S.alpb ? S.keps * S.alpbet / S.adet[m] : 0.0;
  double v;
  if (i == j) {
    const double vs = S.vdwsa[i];
    v = S.keps * (1.0 / S.brad[i]) + 2.0 * (S.hbmag[i] * (S.sasa[i] * (1.0 / (vs * vs))));
  } else {
    const double* xi = &b.a.xyz[3 * i];
    const double* xj = &b.a.xyz[3 * j];
    const double d[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (S.alpb) {
      const double ab = sqrt(S.brad[i] * S.brad[j]);
      double arg = ab / (ab + axtb_solv::ZETA_P16 / 16.0 * r);
      arg *= arg;
      arg *= arg;
      arg *= arg;
      arg *= arg;
      v = S.keps * (1.0 / (r + ab * arg));
    } else {
      const double r2 = r * r, aa = S.brad[i] * S.brad[j], dd = 0.25 * r2 / aa;
      v = S.keps * (1.0 / sqrt(r2 + aa * exp(-dd)));
    }
  }
  S.B[M.pair0 + e] = v + shift;
}

// (B q)_a over the lanes, in a fixed order
__device__ __forceinline__ double solv_bq(const SolvDev& S, const Mol& M, int a, const double* q, int lane) {
  const double* row = S.B + M.pair0 + (long long)(a - M.at0) * M.nat;
  double s = 0.0;
  for (int j = lane; j < M.nat; j += 64) s += row[j] * q[M.at0 + j];
  return wave_sum(s);
}

// potential of the input charges: vat += B q (lib.rs:653-661). One wavefront per atom.
__global__ __launch_bounds__(256) void k_solv_pot(Batch b, const int* active, SolvDev S, const double* q, double* vat) {
  const int a = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (a >= b.nat) return;
  const int m = b.a.mol[a];
  if (active && !active[m]) return;
  const double v = solv_bq(S, b.mol[m], a, q, lane);
  if (lane == 0) vat[a] += v;
}

// energy of the output charges q_a (B q)_a / 2 into the atom's electronic energy and term slot 5
// (lib.rs:640-650)
__global__ __launch_bounds__(256) void k_solv_energy(Batch b, const int* active, SolvDev S, const double* q,
                                                     double* eel, double* eterm) {
  const int a = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (a >= b.nat) return;
  const int m = b.a.mol[a];
  if (active && !active[m]) return;
  const double v = solv_bq(S, b.mol[m], a, q, lane);
  if (lane == 0) {
    const double e = q[a] * (0.5 * v);
    eel[a] += e;
    eterm[8 * (long long)a + 5] = e;
  }
}

// a molecule whose surface could not be built fails its SCF (status bit 32)
__global__ void k_solv_flag(const int* fail, int nmol, int* active, int* status) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < nmol && fail[m] && active[m]) {
    status[m] |= 32;
    active[m] = 0;
  }
}

// Gradient of q^T B q / 2 at fixed charges through the kernel's explicit distance dependence, and
// dE/dbrad (kernel.rs:75-157), times dbrad/dpsi into work. One wavefront per atom: the pair terms
// of atom a with every partner (each pair's derivative is odd in a <-> j).
__global__ __launch_bounds__(256) void k_solv_grad_pairs(Batch b, const int* ok, SolvDev S, const double* q,
                                                         double* grad) {
  const int a = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (a >= b.nat) return;
  const int m = b.a.mol[a];
  if (!ok[m]) return;
  const Mol M = b.mol[m];
  const double* xa = &b.a.xyz[3 * a];
  const double keps = S.keps, ba = S.brad[a];
  double g[3] = {0, 0, 0}, de = 0.0;
  for (int j = M.at0 + lane; j < M.at0 + M.nat; j += 64) {
    if (j == a) continue;
    const double* xj = &b.a.xyz[3 * j];
    const double v[3] = {xa[0] - xj[0], xa[1] - xj[1], xa[2] - xj[2]};
    const double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double qq = q[a] * q[j], bj = S.brad[j];
    if (S.alpb) {
      using axtb_solv::ZETA_P16;
      const double ab = sqrt(ba * bj);
      const double arg1 = ab / (ab + ZETA_P16 / 16.0 * r);
      double arg16 = arg1 * arg1;
      arg16 *= arg16;
      arg16 *= arg16;
      arg16 *= arg16;
      const double fgb = r + ab * arg16, dfgb = 1.0 / fgb, dfgb2 = dfgb * dfgb;
      const double ap = (1.0 - ZETA_P16 * arg1 * arg16) * dfgb2;
      for (int k = 0; k < 3; k++) g[k] -= ap * v[k] * keps / r * qq;
      const double bp = -0.5 * (r * ZETA_P16 / ab * arg1 + 1.0) / ab * arg16 * dfgb2;
      de += bj * bp * keps * qq;
    } else {
      const double r2 = r * r, aa = ba * bj, dd = 0.25 * r2 / aa, expd = exp(-dd);
      const double fgb2 = r2 + aa * expd, dfgb2 = 1.0 / fgb2, dfgb = sqrt(dfgb2);
      const double dfgb3 = dfgb2 * dfgb * keps, ap = (1.0 - 0.25 * expd) * dfgb3;
      for (int k = 0; k < 3; k++) g[k] -= ap * v[k] * qq;
      const double bp = -0.5 * expd * (1.0 + dd) * dfgb3;
      de += bj * bp * qq;
    }
  }
  for (int k = 0; k < 3; k++) g[k] = wave_sum(g[k]);
  de = wave_sum(de);
  if (lane) return;
  const double bp = 1.0 / ba, qb = q[a] * bp;
  de += -0.5 * keps * qb * bp * q[a];
  S.work[a] = de * S.dbrad[a];
  for (int k = 0; k < 3; k++) grad[3 * a + k] += g[k];
}

// sum_i work_i d psi_i / d R (born.rs:162-207): for atom a and partner j, (work_a dpsi_a/dr +
// work_j dpsi_j/dr) along R_a - R_j. One wavefront per atom.
__global__ __launch_bounds__(256) void k_solv_grad_psi(Batch b, const int* ok, SolvDev S, double* grad) {
  const int a = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (a >= b.nat) return;
  const int m = b.a.mol[a];
  if (!ok[m]) return;
  const Mol M = b.mol[m];
  const double* xa = &b.a.xyz[3 * a];
  double g[3] = {0, 0, 0};
  for (int j = M.at0 + lane; j < M.at0 + M.nat; j += 64) {
    if (j == a) continue;
    const double* xj = &b.a.xyz[3 * j];
    const double d[3] = {xa[0] - xj[0], xa[1] - xj[1], xa[2] - xj[2]};
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (!(r2 < S.lrcut2)) continue;
    BornPsi mine, other;
    bool hm, ho;
    solv_born_contrib(S, a, j, sqrt(r2), mine, hm, other, ho);
    double s = 0.0;
    if (hm) s += S.work[a] * mine.dg;
    if (ho) s += S.work[j] * other.dg;
    for (int k = 0; k < 3; k++) g[k] += s * d[k];
  }
  for (int k = 0; k < 3; k++) g[k] = wave_sum(g[k]);
  if (lane == 0)
    for (int k = 0; k < 3; k++) grad[3 * a + k] += g[k];
}

// sum over the grid points p of `from` (weights wsa) of wsa_p dsasa_ij / (sasa_ij |t|) t,
// t = x_p - R_to, over the points where `to` is within its outer smoothing radius: minus
// d sasa_from / d R_to when from != to (the lanes hold partial sums)
__device__ __forceinline__ void solv_point_sum(const Batch& b, const SolvDev& S, int from, int to, int lane,
                                               double acc[3], double scale) {
  const double* xf = &b.a.xyz[3 * from];
  const double* xt = &b.a.xyz[3 * to];
  const double* wsa = S.wsa + (long long)from * S.npts;
  const double rs = S.vdwsa[from], t2b = S.trj2b[to], vst = S.vdwsa[to];
  for (int p = lane; p < S.npts; p += 64) {
    const double ws = wsa[p];
    if (ws == 0.0) continue;
    const double* g = &S.grid[4 * p];
    const double t[3] = {xf[0] + rs * g[0] - xt[0], xf[1] + rs * g[1] - xt[1], xf[2] + rs * g[2] - xt[2]};
    const double tj2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    if (!(tj2 < t2b)) continue;
    const double sqtj = sqrt(tj2), uj = sqtj - vst, ah3uj2 = S.ah3 * uj * uj;
    double dsij = S.ah1 + 3.0 * ah3uj2;
    double sij = S.ah0 + (S.ah1 + ah3uj2) * uj;
    if (1.0 < sij) sij = 1.0;
    if (sij < 0.0) sij = 0.0;
    dsij /= sij * sqtj;
    const double f = scale * ws * dsij;
    for (int k = 0; k < 3; k++) acc[k] += f * t[k];
  }
}

// Surface gradient contracted with c_i = gamma_i + (hbmag_i / R_sa,i^2) q_i^2 (the surface
// energy and the H-bond term, lib.rs:627-636, 700-706; sasa.rs:245-268): for atom j,
// sum over its neighbours n of c_n d sasa_n / d R_j, plus c_j d sasa_j / d R_j. One workgroup of
// 64 per atom; the neighbour lists are the ones of k_solv_sasa (the relation is symmetric).
__global__ __launch_bounds__(64) void k_solv_grad_sasa(Batch b, const int* ok, SolvDev S, const double* q,
                                                       double* grad) {
  __shared__ double nx[3 * SOLV_MAXNB];
  __shared__ int idx[SOLV_MAXNB];
  const int j = blockIdx.x, lane = threadIdx.x;
  const int m = b.a.mol[j];
  if (!ok[m]) return;
  const Mol M = b.mol[m];
  const int nno = solv_neighbours(b, S, j, M, nx, idx);
  if (nno > SOLV_MAXNB) return;  // the molecule has failed already
  auto coef = [&](int a) {
    const double vs = S.vdwsa[a];
    return S.gam[a] + S.hbmag[a] * (1.0 / (vs * vs)) * q[a] * q[a];  // gamma + dhbdw q^2
  };
  const double cj = coef(j);
  double acc[3] = {0, 0, 0};
  for (int sl = 0; sl < nno; sl++) {
    const int n = idx[sl];
    solv_point_sum(b, S, n, j, lane, acc, -coef(n));  // n's surface moved by R_j
    solv_point_sum(b, S, j, n, lane, acc, cj);        // j's own surface against neighbour n
  }
  for (int k = 0; k < 3; k++) acc[k] = wave_sum(acc[k]);
  if (lane == 0)
    for (int k = 0; k < 3; k++) grad[3 * j + k] += acc[k];
}

// Gradient of keps alpbet (sum q)^2 / (2 aDet) (kernel.rs:211-238). One workgroup per molecule.
__global__ __launch_bounds__(256) void k_solv_grad_adet(Batch b, const int* ok, SolvDev S, const double* q,
                                                        double* grad) {
  __shared__ double red[16];
  const int m = blockIdx.x;
  if (!ok[m]) return;
  const Mol M = b.mol[m];
  double i[3][3], c[3], t;
  solv_inertia(b, S, M, red, i, c, t);
  double qs = 0.0;
  for (int a = M.at0 + threadIdx.x; a < M.at0 + M.nat; a += blockDim.x) qs += q[a];
  const double qt = block_sum(qs, red);
  const double adet = sqrt(pow(solv_det3(i), 1.0 / 3.0) / (0.4 * t));
  const double t2 = t * t, a2 = adet * adet;
  const double fa = 250.0 / (48.0 * (t2 * t) * ((a2 * a2) * adet));
  const double fb = -0.5 * (S.keps * S.alpbet) * (qt * qt) / a2;
  const double xy = i[0][1] * i[2][2] - i[0][2] * i[1][2], xz = i[0][2] * i[1][1] - i[0][1] * i[2][1];
  const double yz = i[0][0] * i[1][2] - i[0][1] * i[0][2];
  const double xx = i[0][0] * (i[1][1] + i[2][2]) - i[0][1] * i[0][1] - i[0][2] * i[0][2];
  const double yy = i[1][1] * (i[0][0] + i[2][2]) - i[0][1] * i[0][1] - i[1][2] * i[1][2];
  const double zz = i[2][2] * (i[0][0] + i[1][1]) - i[0][2] * i[0][2] - i[1][2] * i[1][2];
  auto f = [&](double x) { return x * fa * fb; };
  const double d[3][3] = {{f(xx), f(xy), f(xz)}, {f(xy), f(yy), f(yz)}, {f(xz), f(yz), f(zz)}};
  for (int a = M.at0 + threadIdx.x; a < M.at0 + M.nat; a += blockDim.x) {
    const double r = S.vdwr[a], r3 = r * r * r;
    const double v[3] = {b.a.xyz[3 * a] - c[0], b.a.xyz[3 * a + 1] - c[1], b.a.xyz[3 * a + 2] - c[2]};
    for (int k = 0; k < 3; k++) grad[3 * a + k] += r3 * (d[k][0] * v[0] + d[k][1] * v[1] + d[k][2] * v[2]);
  }
}

}  // namespace axtb
