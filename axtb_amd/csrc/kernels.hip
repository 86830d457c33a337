// GFN2-xTB kernels of the HIP engine: geometry setup (coordination numbers, EEQ system, gamma,
// repulsion, D4), the integral build, the SCC iteration pieces (potentials, Fock update, Fermi
// filling, Mulliken moments, energy, Broyden mixing) and the analytic gradient.
//
// Every kernel works on a batch of molecules (Batch/Mol descriptors); a single molecule is a
// batch of one. Per-molecule work is skipped where `active[m] == 0` (converged or failed).
// Reductions run in a fixed order, so results do not depend on scheduling.
#include "common.hpp"
#include "shellpair.hpp"

namespace axtb {

__global__ void k_fill(double* p, long long n, double v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long st = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = v;
}

// out[i] = a[i] * b[i] (mul) or a[i] + b[i]
__global__ void k_pairwise(const double* a, const double* b, long long n, int mul, double* out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long st = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) out[i] = mul ? a[i] * b[i] : a[i] + b[i];
}

// ------------------------------------------------------------------ counting functions
__device__ __forceinline__ double exp_count(double k, double r, double r0) { return 1.0 / (1.0 + exp(-k * (r0 / r - 1.0))); }
__device__ __forceinline__ double exp_dcount(double k, double r, double r0) {
  double e = exp(-k * (r0 / r - 1.0));
  return (-k * r0 * e) / (r * r * (e + 1.0) * (e + 1.0));
}
// GFN2 double-exponential CN (base/src/ncoord.rs:68-115)
__device__ __forceinline__ double dexp_count(double r, double rc) { return exp_count(10.0, r, rc) * exp_count(20.0, r, rc + 2.0); }
__device__ __forceinline__ double dexp_dcount(double r, double rc) {
  return exp_dcount(10.0, r, rc) * exp_count(20.0, r, rc + 2.0) + exp_count(10.0, r, rc) * exp_dcount(20.0, r, rc + 2.0);
}
// dftd4 erf counting (base/src/ncoord.rs:140-152), k = 7.5
__device__ __forceinline__ double erf_count(double r, double rc) { return 0.5 * (1.0 + erf(-((7.5 * (r - rc)) / rc))); }
__device__ __forceinline__ double erf_dcount(double r, double rc) {
  double e = 7.5 * (r - rc) / rc;
  return -(7.5 * exp(-(e * e))) / (1.7724538509055159 * rc);
}
__device__ __forceinline__ double d4_en_factor(double ei, double ej) {  // d4/src/ncoord.rs:19-27
  double x = fabs(ei - ej) + 19.08857;
  return 4.10451 * exp(-((x * x) / (2.0 * (11.28174 * 11.28174))));
}

// ------------------------------------------------------------------ setup: per-atom quantities
// One wavefront per atom: the GFN2 CN (40 bohr, tblite's conventions 25), the EEQ CN (multicharge,
// 25 bohr, capped at 8) and the D4 CN (40 bohr, tblite 30; EN-weighted); then shell levels and AES
// radii of the atom.
__global__ void k_atom_setup(Batch b, const int* active, double* cn, double* cn_eeq, double* cn_d4, double* se,
                             double* dsedcn, double* mrad, double* dmrdcn) {
  int a = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (a >= b.nat) return;
  int m = b.a.mol[a];
  if (active && !active[m]) return;
  const Mol M = b.mol[m];
  const double* xa = &b.a.xyz[3 * a];
  double c1 = 0.0, c2 = 0.0, c3 = 0.0;
  for (int j = M.at0 + lane; j < M.at0 + M.nat; j += 64) {
    if (j == a) continue;
    const double* xj = &b.a.xyz[3 * j];
    double v[3] = {xa[0] - xj[0], xa[1] - xj[1], xa[2] - xj[2]};
    double r2 = dot3(v, v);
    if (r2 < 1e-12) continue;
    double r = sqrt(r2);
    if (r2 <= b.cn_cut2) c1 += dexp_count(r, b.a.rcov_cn[a] + b.a.rcov_cn[j]);
    if (r2 <= b.d4cn_cut2) c3 += d4_en_factor(b.a.en[a], b.a.en[j]) * erf_count(r, b.a.d4_rcov[a] + b.a.d4_rcov[j]);
    if (r2 <= 625.0) c2 += erf_count(r, b.a.rcov_eeq[a] + b.a.rcov_eeq[j]);
  }
  c1 = wave_sum(c1);
  c2 = wave_sum(c2);
  c3 = wave_sum(c3);
  if (lane == 0) {
    cn[a] = c1;
    // log cap at 8 (base/src/ncoord.rs:258-262)
    cn_eeq[a] = log(1.0 + exp(8.0)) - log(exp(8.0 - c2) + 1.0);
    cn_d4[a] = c3;
    // AES radii (coulomb.rs:368-381)
    double arg = c1 - b.a.aes_vcn[a] - 1.2;
    double t1 = exp(-4.0 * arg);
    double t2 = (5.0 - b.a.aes_rad[a]) / (1.0 + t1);
    mrad[a] = b.a.aes_rad[a] + t2;
    dmrdcn[a] = -t2 * 4.0 * t1 / (1.0 + t1);
    // shell levels (hamiltonian.rs:42-54)
    for (int s = b.a.ish[a]; s < b.a.ish[a] + b.a.nsh[a]; s++) {
      se[s] = b.s.se0[s] - b.s.kcn[s] * c1;
      dsedcn[s] = -b.s.kcn[s];
    }
  }
}

// gamma of the shell pair (coulomb.rs:61-119, exponent 2), stored nsh x nsh per molecule
__global__ void k_gamma(Batch b, double* gam) {
  int m = blockIdx.y;
  const Mol M = b.mol[m];
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)M.nsh * M.nsh) return;
  int i = (int)(idx / M.nsh), j = (int)(idx % M.nsh);
  int si = M.sh0 + i, sj = M.sh0 + j;
  int ai = b.s.at[si], aj = b.s.at[sj];
  double eta = 0.5 * (b.s.hard[si] + b.s.hard[sj]);
  double val;
  if (ai == aj) {
    val = eta;
  } else {
    const double* x = &b.a.xyz[3 * ai];
    const double* y = &b.a.xyz[3 * aj];
    double v[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
    val = 1.0 / sqrt(dot3(v, v) + 1.0 / (eta * eta));
  }
  gam[M.shp0 + idx] = val;
}

// EEQ-2019 bordered system (axtb/src/eeq.rs:43-77, base/src/eeq.rs:69-104) of each molecule:
// (nat+1)^2 column-major at eq0 = the molecule's pair0 + its index (pair pool sized for it).
__global__ void k_eeq_system(Batch b, const long long* eq0, const double* cn_eeq, double* A, double* rhs) {
  int m = blockIdx.y;
  const Mol M = b.mol[m];
  int n1 = M.nat + 1;
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n1 * n1) return;
  int i = (int)(idx % n1), j = (int)(idx / n1);
  double v;
  if (i == M.nat && j == M.nat) {
    v = 0.0;
  } else if (i == M.nat || j == M.nat) {
    v = 1.0;
  } else {
    int ai = M.at0 + i, aj = M.at0 + j;
    if (i == j) {
      v = b.a.eeq_eta[ai] + 0.7978845608028654 / b.a.eeq_rad[ai];  // sqrt(2/pi)
    } else {
      const double* x = &b.a.xyz[3 * ai];
      const double* y = &b.a.xyz[3 * aj];
      double d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
      double r2 = dot3(d, d);
      double ri = b.a.eeq_rad[ai], rj = b.a.eeq_rad[aj];
      double gam = 1.0 / (ri * ri + rj * rj);
      v = erf(sqrt(r2 * gam)) / sqrt(r2);
    }
  }
  A[eq0[m] + idx] = v;
  if (j == 0) {
    double r;
    if (i == M.nat) {
      r = M.charge;
    } else {
      int ai = M.at0 + i;
      double c = cn_eeq[ai];
      r = -b.a.eeq_chi[ai] + b.a.eeq_kcnchi[ai] / sqrt(c + 1e-14) * c;
    }
    rhs[(long long)M.at0 + m + i] = r;
  }
}

// q = V diag(1/w) V^T rhs of the EEQ system from its eigendecomposition, then the start shell
// charges split by reference occupation (xtb.rs:530-547). One workgroup per molecule; a
// singular system or non-finite charges leave neutral atoms (the reference's "sad" start).
__global__ void k_eeq_solve(Batch b, const int* mlist, const long long* eq0, const double* V, const double* w,
                            const double* rhs, double* y, double* qstart, int* eeq_failed) {
  __shared__ int bad;
  int m = mlist[blockIdx.x];
  const Mol M = b.mol[m];
  int n1 = M.nat + 1;
  const double* Vm = V + eq0[m];
  const double* r = rhs + (long long)M.at0 + m;
  double* ym = y + (long long)M.at0 + m;
  const double* wm = w + (long long)M.at0 + m;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < n1; k += blockDim.x) {
    double s = 0.0;
    for (int i = 0; i < n1; i++) s += Vm[(long long)k * n1 + i] * r[i];
    if (!(fabs(wm[k]) > 1e-300)) bad = 1;
    ym[k] = s / wm[k];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < M.nat; i += blockDim.x) {
    double q = 0.0;
    for (int k = 0; k < n1; k++) q += Vm[(long long)k * n1 + i] * ym[k];
    if (!isfinite(q)) bad = 1;
    int a = M.at0 + i;
    double tot = 0.0;
    for (int s = b.a.ish[a]; s < b.a.ish[a] + b.a.nsh[a]; s++) tot += b.s.refocc[s];
    for (int s = b.a.ish[a]; s < b.a.ish[a] + b.a.nsh[a]; s++) qstart[s] = tot > 0.0 ? q * b.s.refocc[s] / tot : 0.0;
  }
  __syncthreads();
  if (bad) {
    for (int s = M.sh0 + threadIdx.x; s < M.sh0 + M.nsh; s += blockDim.x) qstart[s] = 0.0;
    if (threadIdx.x == 0) eeq_failed[m] = 1;
  }
}

// EEQ start charges of molecule m from A^-1 r (column 0 of x) and A^-1 1 (column 1), A the N x N
// block of the EEQ system [A 1; 1^T 0] [q; lambda] = [r; Q]: lambda = (sum x1 - Q) / sum x2,
// q = x1 - lambda x2 (the bordered system of eeq.rs:43-77 by its Schur complement), split over the
// shells as k_eeq_solve does. One workgroup; sums in a fixed order.
__global__ void k_eeq_schur(Batch b, int m, const double* x, const double* rhs, double* qstart, int* eeq_failed) {
  __shared__ double red[16];
  const Mol M = b.mol[m];
  const int n = M.nat;
  const double* r = rhs + (long long)M.at0 + m;
  double s1 = 0.0, s2 = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    s1 += x[i];
    s2 += x[n + i];
  }
  s1 = block_sum(s1, red);
  s2 = block_sum(s2, red);
  const double lam = (s1 - r[n]) / s2;
  bool bad = !isfinite(lam);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double q = x[i] - lam * x[n + i];
    const int a = M.at0 + i;
    double tot = 0.0;
    for (int s = b.a.ish[a]; s < b.a.ish[a] + b.a.nsh[a]; s++) tot += b.s.refocc[s];
    for (int s = b.a.ish[a]; s < b.a.ish[a] + b.a.nsh[a]; s++)
      qstart[s] = (tot > 0.0 && !bad) ? q * b.s.refocc[s] / tot : 0.0;
  }
  if (threadIdx.x == 0 && bad) eeq_failed[m] = 1;
}

// Right-hand sides [r | 1] (N x 2) of k_eeq_schur's solves.
__global__ void k_eeq_rhs(Batch b, int m, const double* rhs, double* R) {
  const Mol M = b.mol[m];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M.nat) return;
  R[i] = rhs[(long long)M.at0 + m + i];
  R[M.nat + i] = 1.0;
}

// ------------------------------------------------------------------ integrals
// hscale of shells (sa on atom A, sb on atom B) (gfn2.rs:24-30,76-98)
__device__ __forceinline__ double hscale(const Batch& b, int sa, int sb) {
  const double KD[3] = {1.85, 2.23, 2.23};
  int la = b.s.l[sa], lb = b.s.l[sb];
  // kshell: 2 between a d shell and an s or p shell, else the average of KDIAG (gfn2.rs:24-30)
  double km = ((la == 2 && lb <= 1) || (lb == 2 && la <= 1)) ? 2.0 : (KD[la] + KD[lb]) / 2.0;
  double dEN = b.a.en[b.s.at[sa]] - b.a.en[b.s.at[sb]];
  double za = b.s.zeta[sa], zb = b.s.zeta[sb];
  double zij = sqrt(2.0 * sqrt(za * zb) / (za + zb));
  return zij * (km * (1.0 + 0.02 * (dEN * dEN)));
}

// Ket rows of a shell-pair type computed by one work item: all of them for s and p pairs, one row
// when either shell is d (keeps the accumulators of a work item at most 5 blocks x 10 moments).
template <int LI, int LJ> struct RowSplit {
  static constexpr bool split = LI == 2 || LJ == 2;
  static constexpr int NR = split ? 1 : 2 * LI + 1;   // rows per work item
  static constexpr int NCH = split ? 2 * LI + 1 : 1;  // work items per shell pair
};

// Ket rows [IM0, IM0 + NR) of one shell-pair block of the integral pools: element (mu, nu) and,
// off-site, its transpose with the moments shifted to the bra atom.
template <int LI, int LJ, int IM0, int NR>
__device__ __forceinline__ void integral_block(const Batch& b, const Mol& M, int si, int sj, const double* vec,
                                               bool onsite, double hij, double* pool, long long stride) {
  constexpr int NSJ = 2 * LJ + 1;
  double blk[NR * NSJ][10];
  shellpair_multipole_t<LI, LJ, IM0, NR>(b.s.nprim[si], &b.s.alpha[MAXPRIM * si], &b.s.coeff[MAXPRIM * si],
                                         b.s.nprim[sj], &b.s.alpha[MAXPRIM * sj], &b.s.coeff[MAXPRIM * sj], vec,
                                         b.intcut, blk);
  const int ii = b.s.iao[si], jj = b.s.iao[sj];
  const long long n = M.nao;
#pragma unroll
  for (int r = 0; r < NR; r++)
#pragma unroll
    for (int jm = 0; jm < NSJ; jm++) {
      const double* o = blk[jm + NSJ * r];
      long long mu = jj + jm, nu = ii + IM0 + r;
      long long e1 = M.mat0 + mu + nu * n;
      pool[e1] = o[0];
      pool[stride + e1] = o[0] * hij;
#pragma unroll
      for (int k = 0; k < 3; k++) pool[(2 + k) * stride + e1] = o[1 + k];
#pragma unroll
      for (int k = 0; k < 6; k++) pool[(5 + k) * stride + e1] = o[4 + k];
      if (!onsite) {
        double dj[3], qj[6];
        shift_to_bra(vec, o[0], &o[1], &o[4], dj, qj);
        long long e2 = M.mat0 + nu + mu * n;
        pool[e2] = o[0];
        pool[stride + e2] = o[0] * hij;
#pragma unroll
        for (int k = 0; k < 3; k++) pool[(2 + k) * stride + e2] = dj[k];
#pragma unroll
        for (int k = 0; k < 6; k++) pool[(5 + k) * stride + e2] = qj[k];
      }
    }
}

// integral_block of the row chunk `ch` (split types: row ch; others: all rows)
template <int LI, int LJ, int IM = 0>
__device__ __forceinline__ void integral_rows(int ch, const Batch& b, const Mol& M, int si, int sj, const double* vec,
                                              bool onsite, double hij, double* pool, long long stride) {
  using RS = RowSplit<LI, LJ>;
  if constexpr (!RS::split) {
    integral_block<LI, LJ, 0, RS::NR>(b, M, si, sj, vec, onsite, hij, pool, stride);
  } else if constexpr (IM < RS::NCH) {
    if (ch == IM)
      integral_block<LI, LJ, IM, 1>(b, M, si, sj, vec, onsite, hij, pool, stride);
    else
      integral_rows<LI, LJ, IM + 1>(ch, b, M, si, sj, vec, onsite, hij, pool, stride);
  }
}

// S, H0, D (3), Q (6) of every shell pair (ish of atom A >= jsh of atom B): each AO element
// written once (ao.rs:195-291, hamiltonian.rs:77-94). Pools are `stride` elements apart:
// slot 0 S, 1 H0, 2-4 D, 5-10 Q; element (mu, nu) of molecule m at mat0 + mu + nu nao; elements of
// screened shell pairs stay as they are (zero). Split by shell-pair type, so that a wavefront runs
// one type: k_int_keys lists the shell pairs (i, j) of each key 3 l_i + l_j that pass the screening
// (entry (m, i nsh + j); the lists grouped by molecule like k_h0_keys' lists, capacity of a group the
// sum of its nsh^2), k_integrals_key<LI, LJ> (grid.y = group, grid.z = row chunk of the types with a
// d shell) computes them.
constexpr int INT_GROUPS = 64;
constexpr int NKEYS = 9;  // shell-pair types (l_i, l_j), l <= 2
__global__ __launch_bounds__(64) void k_int_keys(Batch b, int* cnt, int2* list, long long cap, const long long* goff) {
  const int m = blockIdx.y;
  const Mol M = b.mol[m];
  const long long idx = (long long)blockIdx.x * 64 + threadIdx.x;
  int key = -1;
  if (idx < (long long)M.nsh * M.nsh) {
    const int i = (int)(idx / M.nsh), j = (int)(idx % M.nsh);  // ket shell i (column), bra shell j (row)
    const int si = M.sh0 + i, sj = M.sh0 + j;
    const int ai = b.s.at[si], aj = b.s.at[sj];
    if (ai >= aj) {
      bool keep = true;
      if (ai != aj) {
        // atom-pair and shell-pair exponent screening (integrals.rs:216-218)
        const double* xi = &b.a.xyz[3 * ai];
        const double* xj = &b.a.xyz[3 * aj];
        const double vec[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
        const double r2 = dot3(vec, vec);
        const double a1 = b.s.minalpha[si], a2 = b.s.minalpha[sj];
        keep = !(a1 * a2 * r2 / (a1 + a2) > b.intcut * (1.0 + 1e-10)) && !(r2 > b.mol[m].pair_cut2);
      }
      if (keep) key = 3 * b.s.l[si] + b.s.l[sj];
    }
  }
  const int lane = threadIdx.x, g = m % INT_GROUPS;
  for (int k = 0; k < NKEYS; k++) {
    const bool has = key == k;
    const unsigned long long bal = __ballot(has);
    if (!bal) continue;
    int base = 0;
    if (lane == 0) base = atomicAdd(&cnt[k * INT_GROUPS + g], (int)__popcll(bal));
    base = __shfl(base, 0, 64);
    if (has) list[k * cap + goff[g] + base + __popcll(bal & ((1ull << lane) - 1ull))] = make_int2(m, (int)idx);
  }
}

template <int LI, int LJ>
__global__ __launch_bounds__(64) void k_integrals_key(Batch b, const int2* list, const int* count,
                                                      const long long* goff, const double* se, double* pool,
                                                      long long stride) {
  const int g = blockIdx.y, t = blockIdx.x * 64 + threadIdx.x;
  if (t >= count[g]) return;
  const int2 it = list[goff[g] + t];
  const Mol M = b.mol[it.x];
  const int i = it.y / M.nsh, j = it.y % M.nsh;
  const int si = M.sh0 + i, sj = M.sh0 + j;
  const int ai = b.s.at[si], aj = b.s.at[sj];
  const double* xi = &b.a.xyz[3 * ai];
  const double* xj = &b.a.xyz[3 * aj];
  double vec[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
  const double r2 = dot3(vec, vec);
  const bool onsite = ai == aj;
  double hij;
  if (onsite) {
    hij = 0.5 * (se[si] + se[sj]);
  } else {
    const double rr = sqrt(sqrt(r2) / (b.a.rad[aj] + b.a.rad[ai]));
    const double sp = (1.0 + b.s.shpoly[si] * rr) * (1.0 + b.s.shpoly[sj] * rr);
    hij = 0.5 * (se[si] + se[sj]) * hscale(b, sj, si) * sp;
  }
  integral_rows<LI, LJ>(blockIdx.z, b, M, si, sj, vec, onsite, hij, pool, stride);
}
#define AXTB_INT_KEY(LI, LJ)                                                                                          \
  template __global__ void k_integrals_key<LI, LJ>(Batch, const int2*, const int*, const long long*, const double*, \
                                                   double*, long long);
AXTB_INT_KEY(0, 0)
AXTB_INT_KEY(0, 1)
AXTB_INT_KEY(0, 2)
AXTB_INT_KEY(1, 0)
AXTB_INT_KEY(1, 1)
AXTB_INT_KEY(1, 2)
AXTB_INT_KEY(2, 0)
AXTB_INT_KEY(2, 1)
AXTB_INT_KEY(2, 2)
#undef AXTB_INT_KEY

// ------------------------------------------------------------------ repulsion (repulsion.rs:44-73)
// One wavefront per atom: E_A (half of each pair) and dE/dR_A, cutoff 40 bohr (tblite 25).
__global__ void k_repulsion(Batch b, const int* active, double* e, double* g) {
  int a = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (a >= b.nat) return;
  int m = b.a.mol[a];
  if (active && !active[m]) return;
  const Mol M = b.mol[m];
  const double* xa = &b.a.xyz[3 * a];
  double es = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
  for (int j = M.at0 + lane; j < M.at0 + M.nat; j += 64) {
    if (j == a) continue;
    const double* xj = &b.a.xyz[3 * j];
    double v[3] = {xa[0] - xj[0], xa[1] - xj[1], xa[2] - xj[2]};
    double r2 = dot3(v, v);
    if (r2 > b.rep_cut2 || r2 < 1e-12) continue;
    double r1 = sqrt(r2);
    double kexp = (b.a.z[a] > 2 || b.a.z[j] > 2) ? 1.5 : 1.0;
    double alpha = sqrt(b.a.rep_alpha[j] * b.a.rep_alpha[a]);
    double zeff = b.a.rep_zeff[j] * b.a.rep_zeff[a];
    double r1k = kexp == 1.5 ? r1 * sqrt(r1) : r1;
    double de = zeff * exp(-alpha * r1k) / r1;
    es += 0.5 * de;
    double f = -(alpha * r1k * kexp + 1.0) * de / r2;
    gx += f * v[0];
    gy += f * v[1];
    gz += f * v[2];
  }
  es = wave_sum(es);
  gx = wave_sum(gx);
  gy = wave_sum(gy);
  gz = wave_sum(gz);
  if (lane == 0) {
    e[a] = es;
    g[3 * a] = gx;
    g[3 * a + 1] = gy;
    g[3 * a + 2] = gz;
  }
}

// ------------------------------------------------------------------ D4 (crates/d4/src)
__device__ __forceinline__ void zeta_dz(double a, double c, double qref, double qmod, double& z, double& dz) {
  if (qmod < 0.0) {
    z = exp(a);
    dz = 0.0;
    return;
  }
  double e = exp(c * (1.0 - qref / qmod));
  z = exp(a * (1.0 - e));
  dz = -((a * c * e * z * qref) / (qmod * qmod));
}

// Reference weights of every atom (d4/src/model.rs:468-545) at charges q: gw, and either
// dgw/dq (mode 0) or dgw/dCN (mode 1) into dgw. One thread per atom.
__global__ void k_d4_weights(Batch b, const int* active, const double* cn_d4, const double* q, double* gw,
                             double* dgw, int mode) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= b.nat) return;
  if (active && !active[b.a.mol[a]]) return;
  const int nref = b.a.d4_nref[a];
  const double* refcn = &b.a.d4_refcn[MAXREF * a];
  const int* ngw = &b.a.d4_ngw[MAXREF * a];
  double cni = cn_d4[a];
  double norm = 0.0, dnorm = 0.0;
  for (int r = 0; r < nref; r++)
    for (int g = 1; g <= ngw[r]; g++) {
      double wf = g * 6.0, d = cni - refcn[r];
      double gg = exp(-(wf * (d * d)));
      norm += gg;
      dnorm += 2.0 * wf * (refcn[r] - cni) * gg;
    }
  norm = fabs(norm) > 1.4916681462400413e-154 ? 1.0 / norm : 0.0;
  double maxcn = -INFINITY;
  for (int r = 0; r < nref; r++) maxcn = fmax(maxcn, refcn[r]);
  for (int r = 0; r < MAXREF; r++) {
    double w = 0.0, dw = 0.0;
    if (r < nref) {
      double expw = 0.0, expd = 0.0;
      for (int g = 1; g <= ngw[r]; g++) {
        double wf = g * 6.0, d = cni - refcn[r];
        double gg = exp(-(wf * (d * d)));
        expw += gg;
        expd += 2.0 * wf * (refcn[r] - cni) * gg;
      }
      double gwk = expw * norm;
      if (!isfinite(gwk) || norm == 0.0) gwk = fabs(maxcn - refcn[r]) < 1e-12 ? 1.0 : 0.0;
      double z, dz;
      zeta_dz(3.0, b.a.d4_gi[a], b.a.d4_refq[MAXREF * a + r] + b.a.d4_zeff[a], q[a] + b.a.d4_zeff[a], z, dz);
      w = gwk * z;
      if (mode == 0) {
        dw = gwk * dz;
      } else {
        double dgwk = norm * (expd - expw * dnorm * norm);
        if (!isfinite(dgwk) || norm == 0.0) dgwk = 0.0;
        dw = dgwk * z;
      }
    }
    gw[MAXREF * a + r] = w;
    dgw[MAXREF * a + r] = dw;
  }
}

// C6 of a pair and its derivative contractions (model.rs:607-627)
__device__ __forceinline__ void pair_c6(const Batch& b, int i, int j, const double* gw, const double* dgw, double& c6,
                                        double& di, double& dj) {
  // the references of j over a fixed 7 (every load in bounds, unused ones masked to zero), so
  // that a row's loads are issued together; the sums run in the order of the ni x nj loops (the
  // added zeros leave them unchanged)
  const double* tab = &b.c6ref[((long long)b.a.sp[i] * b.nsp + b.a.sp[j]) * 49];
  const int ni = b.a.d4_nref[i], nj = b.a.d4_nref[j];
  double wj[MAXREF], dwj[MAXREF];
#pragma unroll
  for (int r = 0; r < MAXREF; r++) {
    wj[r] = r < nj ? gw[MAXREF * j + r] : 0.0;
    dwj[r] = r < nj ? dgw[MAXREF * j + r] : 0.0;
  }
  c6 = di = dj = 0.0;
  for (int ri = 0; ri < ni; ri++) {
    const double wi = gw[MAXREF * i + ri], dwi = dgw[MAXREF * i + ri];
#pragma unroll
    for (int rj = 0; rj < MAXREF; rj++) {
      const double r = rj < nj ? tab[ri * 7 + rj] : 0.0;
      c6 += wi * wj[rj] * r;
      di += dwi * wj[rj] * r;
      dj += wi * dwj[rj] * r;
    }
  }
}

__device__ __forceinline__ double pw3(double x) { return (x * x) * x; }
__device__ __forceinline__ double pw4(double x) { double x2 = x * x; return x2 * x2; }

// d4/src/disp.rs:43-63: 5th-order smoothstep over the last `width` bohr
__device__ __forceinline__ void smooth_cut(double r, double cut, double width, double& sw, double& dsw) {
  double w = fmin(width, cut), inner = cut - w;
  if (r <= inner) {
    sw = 1.0;
    dsw = 0.0;
  } else if (r >= cut) {
    sw = 0.0;
    dsw = 0.0;
  } else {
    double x = (cut - r) / w, x2 = x * x;
    sw = (x2 * x) * (10.0 + x * (-15.0 + 6.0 * x));
    double omx = 1.0 - x;
    dsw = -((30.0 * x2 * (omx * omx)) / w);
  }
}

// rational-damped pair energy factor with the smooth two-body cutoff (axtb 72 bohr switched over
// the last 1 bohr, tblite 50 / 0.05; disp.rs:240-272)
__device__ __forceinline__ bool d4_pair_e(const Batch& b, int i, int j, const double* v, double r2, double& e,
                                          double& gdr) {
  if (r2 > b.d4_rc_2b * b.d4_rc_2b || r2 < 2.220446049250313e-16) return false;
  double r = sqrt(r2), sw, dsw;
  smooth_cut(r, b.d4_rc_2b, b.d4_w_2b, sw, dsw);
  if (sw <= 0.0) return false;
  double rr = 3.0 * b.a.d4_r4r2[i] * b.a.d4_r4r2[j];
  double r0 = 0.52 * sqrt(rr) + 5.0;
  double r06 = pw3(r0 * r0);
  double t6 = 1.0 / (pw3(r2) + r06), t8 = 1.0 / (pw4(r2) + r06 * r0 * r0);
  double d6 = -(6.0 * (r2 * r2) * (t6 * t6)), d8 = -(8.0 * pw3(r2) * (t8 * t8));
  double e0 = t6 + 2.7 * rr * t8, g0 = d6 + 2.7 * rr * d8;
  e = sw * e0;
  gdr = sw * g0 + dsw * e0 / r;
  return true;
}

// ATM pair table (disp.rs:548-605): per ordered pair (a, k) of a molecule, nat x nat row-major at
// pair0: r2, 1/r, t = (r0/r)^(alpha/3), w = sqrt|C6| sw; plus C6, dlnC6/dCN_a, sw, dsw/r.
__global__ void k_atm_pairs(Batch b, const double* gw, const double* dgw, double* P) {
  int m = blockIdx.y;
  const Mol M = b.mol[m];
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long nn = (long long)M.nat * M.nat;
  if (idx >= nn) return;
  int a = M.at0 + (int)(idx / M.nat), k = M.at0 + (int)(idx % M.nat);
  long long o = M.pair0 + idx;
  long long tot = b.mol[b.nmol - 1].pair0 + (long long)b.mol[b.nmol - 1].nat * b.mol[b.nmol - 1].nat;
  double r2v = 0, irv = 0, tv = 0, wv = 0, c6v = 0, qav = 0, swv = 0, dswv = 0;
  if (a != k) {
    const double* xa = &b.a.xyz[3 * a];
    const double* xk = &b.a.xyz[3 * k];
    double v[3] = {xk[0] - xa[0], xk[1] - xa[1], xk[2] - xa[2]};
    double r2 = dot3(v, v);
    if (r2 <= b.d4_rc_3b * b.d4_rc_3b && r2 >= 2.220446049250313e-16) {
      double r = sqrt(r2), sw, dsw;
      smooth_cut(r, b.d4_rc_3b, b.d4_w_3b, sw, dsw);
      double rr = 3.0 * b.a.d4_r4r2[a] * b.a.d4_r4r2[k];
      double r0 = 0.52 * sqrt(rr) + 5.0;
      double c6, da, dk;
      pair_c6(b, a, k, gw, dgw, c6, da, dk);
      r2v = r2;
      irv = 1.0 / r;
      tv = pow(r0 / r, 16.0 / 3.0);
      wv = sqrt(fabs(c6)) * sw;
      c6v = c6;
      qav = c6 != 0.0 ? da / c6 : 0.0;
      swv = sw;
      dswv = dsw / r;
    }
  }
  P[o] = r2v;
  P[tot + o] = irv;
  P[2 * tot + o] = tv;
  P[3 * tot + o] = wv;
  P[4 * tot + o] = c6v;
  P[5 * tot + o] = qav;
  P[6 * tot + o] = swv;
  P[7 * tot + o] = dswv;
}

// ATM three-body energy and gradient by pair sums (disp.rs:607-720, axtb's d4_atm_pair_sums), in
// two passes. k_atm_pairsum_tiled: the pair's (g, d) summed over third atoms k for every unordered
// atom pair (a, b), a < b, written for both orders: for every k, the damped angular term and its
// derivative along r_ab (the triple sum of the pair is symmetric in a and b, so each pair is summed
// once). GD[o] = cab * sum_k (derivative part), GD[tot + o] = cab * sum_k (energy part),
// cab = -5 sqrt|C6_ab|.
// Over 16 x 16 tiles of atom pairs (tile (I, J), J >= I, of molecule m: int4 {m, I, J, 0}), the
// third atoms k in LDS-staged chunks of 32: the rows of the tile's 16 + 16 atoms are read once per
// tile instead of once per pair. One pair (a, b), a < b, per thread.
constexpr int ATM_T = 16, ATM_K = 32;
__global__ __launch_bounds__(256) void k_atm_pairsum_tiled(Batch b, const int4* tiles, const double* P, double* GD) {
  __shared__ double sa[4][ATM_T][ATM_K + 1], sb[4][ATM_T][ATM_K + 1];  // W, T, IR, R2 rows of the tile's atoms
  const int4 tl = tiles[blockIdx.x];
  const Mol M = b.mol[tl.x];
  const long long tot = b.mol[b.nmol - 1].pair0 + (long long)b.mol[b.nmol - 1].nat * b.mol[b.nmol - 1].nat;
  const double *R2 = P, *IR = P + tot, *T = P + 2 * tot, *W = P + 3 * tot, *C6 = P + 4 * tot, *SW = P + 6 * tot;
  const int tid = threadIdx.x, ta = tid >> 4, tb = tid & 15;
  const int la = tl.y * ATM_T + ta, lb = tl.z * ATM_T + tb;
  const bool valid = la < M.nat && lb < M.nat && la < lb;
  const long long rowa = M.pair0 + (long long)la * M.nat, rowb = M.pair0 + (long long)lb * M.nat;
  double swab = 0.0, r2 = 0.0, ir = 0.0, t6 = 0.0;
  if (valid) {
    swab = SW[rowa + lb];
    r2 = R2[rowa + lb];
    ir = IR[rowa + lb];
    t6 = 6.0 * T[rowa + lb];
  }
  const bool act = valid && swab > 0.0;
  const double r2_2 = r2 * r2, alp3 = 16.0 / 3.0;
  const double* src[4] = {W, T, IR, R2};
  double gs = 0.0, ds = 0.0;
  for (int k0 = 0; k0 < M.nat; k0 += ATM_K) {
    __syncthreads();
    // 4 arrays x 16 rows x 32 k for each side: 8 values per thread and side
    for (int e = tid; e < 4 * ATM_T * ATM_K; e += 256) {
      const int arr = e / (ATM_T * ATM_K), r = (e / ATM_K) % ATM_T, k = e % ATM_K;
      const int gk = k0 + k, ra = tl.y * ATM_T + r, rb = tl.z * ATM_T + r;
      sa[arr][r][k] = (gk < M.nat && ra < M.nat) ? src[arr][M.pair0 + (long long)ra * M.nat + gk] : 0.0;
      sb[arr][r][k] = (gk < M.nat && rb < M.nat) ? src[arr][M.pair0 + (long long)rb * M.nat + gk] : 0.0;
    }
    __syncthreads();
    if (!act) continue;
    const int kn = min(ATM_K, M.nat - k0);
    for (int k = 0; k < kn; k++) {
      const double c = sa[0][ta][k] * sb[0][tb][k];
      const double cralp = t6 * sa[1][ta][k] * sb[1][tb][k];
      const double fdmp = 1.0 / (1.0 + cralp);
      const double ir1 = ir * sa[2][ta][k] * sb[2][tb][k];
      const double rrr2 = ir1 * ir1, rrr3 = rrr2 * ir1;
      const double angf = 0.375 * rrr2 * rrr3;
      const double r2ak = sa[3][ta][k], r2bk = sb[3][tb][k];
      const double s1 = r2ak + r2bk, d1 = (r2bk - r2ak) * (r2bk - r2ak);
      const double ang = angf * ((r2_2 - d1) * (s1 - r2)) + rrr3;
      const double cf = c * fdmp;
      ds += ang * cf;
      const double pp = r2 * (r2 * (r2 + s1) + 2.0 * (s1 * s1) + d1) - 5.0 * (d1 * s1);
      gs += cf * (angf * pp - alp3 * (cralp * fdmp) * ang);
    }
  }
  if (!valid) return;
  // (k_atm reads only the pairs with sw > 0)
  const double cab = act ? -(5.0 * sqrt(fabs(C6[rowa + lb]))) : 0.0;
  GD[rowa + lb] = gs * cab;
  GD[rowb + la] = gs * cab;
  GD[tot + rowa + lb] = ds * cab;
  GD[tot + rowb + la] = ds * cab;
}

// k_atm: each atom a keeps its share of its pairs with sw > 0: E_a -= sw d / 6 per pair, the pair
// force along R_b - R_a and dE/dCN_a (writes e3[a], g3[3a], dcn3[a]); one wavefront per atom,
// lanes over the partners b.
__global__ void k_atm(Batch b, const double* P, const double* GD, double* e3, double* g3, double* dcn3) {
  int a = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (a >= b.nat) return;
  int m = b.a.mol[a];
  const Mol M = b.mol[m];
  long long tot = b.mol[b.nmol - 1].pair0 + (long long)b.mol[b.nmol - 1].nat * b.mol[b.nmol - 1].nat;
  const double *R2 = P, *QA = P + 5 * tot, *SW = P + 6 * tot, *DSW = P + 7 * tot;
  int la = a - M.at0;
  long long rowa = M.pair0 + (long long)la * M.nat;
  double es = 0.0, gx = 0.0, gy = 0.0, gz = 0.0, dc = 0.0;
  const double* xa = &b.a.xyz[3 * a];
  for (int lb = lane; lb < M.nat; lb += 64) {
    if (lb == la) continue;
    double swab = SW[rowa + lb];
    if (!(swab > 0.0)) continue;
    const double g = GD[rowa + lb], d = GD[tot + rowa + lb];
    double de = swab * d;
    es -= de / 6.0;
    // force on a from the pair (a, b): -(f) (R_b - R_a) with f = sw g / r2 - dsw/r d
    double f = swab * g / R2[rowa + lb] - DSW[rowa + lb] * d;
    const double* xb = &b.a.xyz[3 * (M.at0 + lb)];
    gx -= f * (xb[0] - xa[0]);
    gy -= f * (xb[1] - xa[1]);
    gz -= f * (xb[2] - xa[2]);
    dc -= 0.5 * de * QA[rowa + lb];
  }
  es = wave_sum(es);
  gx = wave_sum(gx);
  gy = wave_sum(gy);
  gz = wave_sum(gz);
  dc = wave_sum(dc);
  if (lane == 0) {
    e3[a] = es;
    g3[3 * a] = gx;
    g3[3 * a + 1] = gy;
    g3[3 * a + 2] = gz;
    dcn3[a] = dc;
  }
}

// ------------------------------------------------------------------ SCC iteration
// Unpack the packed SCC vector [q_sh | mu (3N) | theta (6N), off-diagonals x sqrt2] of each active
// molecule into shell charges, atomic charges, dipoles and quadrupoles (scf/variables.rs).
__global__ void k_unpack(Batch b, const int* active, const double* v, double* qsh, double* qat, double* dp, double* qp) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= b.nat) return;
  int m = b.a.mol[a];
  if (active && !active[m]) return;
  const Mol M = b.mol[m];
  const double* vm = v + M.vec0;
  int la = a - M.at0;
  double q = 0.0;
  for (int s = b.a.ish[a]; s < b.a.ish[a] + b.a.nsh[a]; s++) {
    double x = vm[s - M.sh0];
    qsh[s] = x;
    q += x;
  }
  qat[a] = q;
  const double inv = 1.0 / 1.4142135623730951;
  for (int c = 0; c < 3; c++) dp[3 * a + c] = vm[M.nsh + 3 * la + c];
  for (int c = 0; c < 6; c++) {
    double x = vm[M.nsh + 3 * M.nat + 6 * la + c];
    qp[6 * a + c] = (c == 1 || c == 3 || c == 4) ? x * inv : x;
  }
}

// Pack moments into the SCC vector layout; with `resid`, also dq = packed - q_in.
__global__ void k_pack(Batch b, const int* active, const double* qsh, const double* dp, const double* qp, double* v,
                       const double* qin, double* dq) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= b.nat) return;
  int m = b.a.mol[a];
  if (active && !active[m]) return;
  const Mol M = b.mol[m];
  int la = a - M.at0;
  const double sq2 = 1.4142135623730951;
  auto put = [&](int k, double x) {
    long long o = M.vec0 + k;
    v[o] = x;
    if (dq) dq[o] = x - qin[o];
  };
  for (int s = b.a.ish[a]; s < b.a.ish[a] + b.a.nsh[a]; s++) put(s - M.sh0, qsh[s]);
  for (int c = 0; c < 3; c++) put(M.nsh + 3 * la + c, dp[3 * a + c]);
  for (int c = 0; c < 6; c++) put(M.nsh + 3 * M.nat + 6 * la + c, qp[6 * a + c] * ((c == 1 || c == 3 || c == 4) ? sq2 : 1.0));
}

// A given SCF start (xtb.rs:519-529, the moments of a previous geometry): the shell charges,
// atomic dipoles and quadrupoles of the molecules with gflag set replace their EEQ or neutral
// start. One thread per atom.
__global__ void k_start_guess(Batch b, const int* gflag, const double* gq, const double* gdp, const double* gqp,
                              double* qsh, double* dp, double* qp) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= b.nat || !gflag[b.a.mol[a]]) return;
  for (int s = b.a.ish[a]; s < b.a.ish[a] + b.a.nsh[a]; s++) qsh[s] = gq[s];
  for (int k = 0; k < 3; k++) dp[3 * a + k] = gdp[3 * a + k];
  for (int k = 0; k < 6; k++) qp[6 * a + k] = gqp[6 * a + k];
}

// v_sh = gamma q (+ Gamma q^2 with `third`): one wavefront per shell row (coulomb.rs:1006-1021,
// 1141-1143).
__global__ void k_shell_pot(Batch b, const int* active, const double* gam, const double* qsh, double* out, int third) {
  int s = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (s >= b.nsh) return;
  int m = b.a.mol[b.s.at[s]];
  if (active && !active[m]) return;
  const Mol M = b.mol[m];
  const double* row = gam + M.shp0 + (long long)(s - M.sh0) * M.nsh;
  double acc = 0.0;
  for (int j = lane; j < M.nsh; j += 64) acc += row[j] * qsh[M.sh0 + j];
  acc = wave_sum(acc);
  if (lane == 0) out[s] = third ? acc + qsh[s] * qsh[s] * b.s.gam3[s] : acc;
}

__constant__ double QSCALE[6] = {1.0, 2.0, 1.0, 2.0, 2.0, 1.0};

// g3, g5 of the damped multipole kernels of a pair (coulomb.rs:383-394)
__device__ __forceinline__ void aes_g(double r1, double mri, double mrj, double& g3, double& g5) {
  double g1 = 1.0 / r1;
  double g3v = g1 * g1 * g1, g5v = g3v * g1 * g1;
  double rr = 0.5 * (mrj + mri) * g1;
  double rr3 = rr * rr * rr;
  g3 = g3v / (1.0 + 6.0 * rr3);
  g5 = g5v / (1.0 + 6.0 * rr3 * rr);
}

// Atom-resolved potentials of one wavefront per atom: AES (coulomb.rs:407-502) into vat, vdp,
// vqp and the D4 dE/dq (disp.rs:420-455) into vat.
__global__ __launch_bounds__(256) void k_atom_pot(Batch b, const int* active, const double* mrad, const double* qat, const double* dp,
                           const double* qp, const double* gw, const double* dgwdq, double* vat, double* vdp,
                           double* vqp) {
  int a = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (a >= b.nat) return;
  int m = b.a.mol[a];
  if (active && !active[m]) return;
  const Mol M = b.mol[m];
  const double* xa = &b.a.xyz[3 * a];
  double va = 0.0, vd[3] = {0, 0, 0}, vq[6] = {0, 0, 0, 0, 0, 0};
  for (int j = M.at0 + lane; j < M.at0 + M.nat; j += 64) {
    if (j == a) continue;
    const double* xj = &b.a.xyz[3 * j];
    double w[3] = {xa[0] - xj[0], xa[1] - xj[1], xa[2] - xj[2]};
    double r2 = dot3(w, w);
    double r1 = sqrt(r2);
    double g3, g5;
    aes_g(r1, mrad[a], mrad[j], g3, g5);
    double tc[6] = {w[0] * w[0], w[0] * w[1], w[1] * w[1], w[0] * w[2], w[1] * w[2], w[2] * w[2]};
    const double* mb = &dp[3 * j];
    double mbw = dot3(mb, w);
    va += mbw * g3;
    for (int k = 0; k < 3; k++) vd[k] += -w[k] * g3 * qat[j] + g5 * (r2 * mb[k] - 3.0 * w[k] * mbw);
    double thb = 0.0;
    for (int k = 0; k < 6; k++) {
      double t = QSCALE[k] * tc[k] * g5;
      vq[k] += t * qat[j];
      thb += t * qp[6 * j + k];
    }
    va += thb;
    double pe, pg;
    if (d4_pair_e(b, a, j, w, r2, pe, pg)) {
      double c6, da, dj;
      pair_c6(b, a, j, gw, dgwdq, c6, da, dj);
      va -= da * pe;
    }
  }
  va = wave_sum(va);
  for (int k = 0; k < 3; k++) vd[k] = wave_sum(vd[k]);
  for (int k = 0; k < 6; k++) vq[k] = wave_sum(vq[k]);
  if (lane == 0) {
    vat[a] = va;
    for (int k = 0; k < 3; k++) vdp[3 * a + k] = vd[k] + 2.0 * b.a.dkernel[a] * dp[3 * a + k];
    for (int k = 0; k < 6; k++) vqp[6 * a + k] = vq[k] + 2.0 * b.a.qkernel[a] * qp[6 * a + k] * QSCALE[k];
  }
}

// Fock matrix F = H0 - 1/2 (B + B^T), B(mu,nu) = S v(nu) + D.vd(A(nu)) + Q.vq(A(nu))
// (scf/fock.rs:22-104). B(nu,mu) is rebuilt from the (mu,nu) integrals by shifting their moments
// to the atom of mu, so every read is coalesced. `out` has the layout of the integral pools.
// F is symmetric, and F(mu, nu) is formed from the (mu, nu) integrals alone (the moments of (nu, mu)
// are those of (mu, nu) shifted to the atom of mu), so only the lower triangle of the integral pools
// is read: one workgroup per 32 x 32 tile (I, J), I >= J, of a molecule (tiles[t] = (m, I << 16 | J)),
// each thread four elements; the tile is written as it is and, through LDS, as its transpose
// (both coalesced). In a diagonal tile the elements above the diagonal come from the transpose.
__global__ __launch_bounds__(256) void k_fock(Batch b, const int* active, const int2* tiles, const double* pool,
                                              long long stride, const double* vsh, const double* vat,
                                              const double* vdp, const double* vqp, double* F) {
  __shared__ double tt[32][33];
  // per row (side 0, mu) and column (side 1, nu) of the tile: the AO potential, the atom's dipole
  // and quadrupole potentials and position (13 values), gathered once per tile
  __shared__ double side[2][32][14];
  const int2 tl = tiles[blockIdx.x];
  const int m = tl.x;
  if (active && !active[m]) return;
  const Mol M = b.mol[m];
  const int n = M.nao, I = tl.y >> 16, J = tl.y & 0xffff;
  const int r = threadIdx.x & 31, c0 = threadIdx.x >> 5;
  // the tile's integrals: all 44 loads of this thread issued before any use
  double s4[4], h4[4], d4[4][3], q4[4][6];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int c = c0 + 8 * k;
    const int mu = 32 * I + r, nu = 32 * J + c;
    const bool in = mu < n && nu < n && mu >= nu;
    const long long e = M.mat0 + (in ? mu + (long long)nu * n : 0);
    s4[k] = in ? pool[e] : 0.0;
    h4[k] = in ? pool[stride + e] : 0.0;
#pragma unroll
    for (int x = 0; x < 3; x++) d4[k][x] = in ? pool[(2 + x) * stride + e] : 0.0;
#pragma unroll
    for (int x = 0; x < 6; x++) q4[k][x] = in ? pool[(5 + x) * stride + e] : 0.0;
  }
  if (threadIdx.x < 64) {
    const int sd = threadIdx.x >> 5, i = threadIdx.x & 31;
    const int ao = 32 * (sd == 0 ? I : J) + i;
    double* o = side[sd][i];
    if (ao < n) {
      const int g = M.ao0 + ao, A = b.ao_at[g];
      o[0] = vsh[b.ao_sh[g]] + vat[A];
      for (int x = 0; x < 3; x++) o[1 + x] = vdp[3 * A + x];
      for (int x = 0; x < 6; x++) o[4 + x] = vqp[6 * A + x];
      for (int x = 0; x < 3; x++) o[10 + x] = b.a.xyz[3 * A + x];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int c = c0 + 8 * k;
    const int mu = 32 * I + r, nu = 32 * J + c;
    double f = 0.0;
    if (mu < n && nu < n && mu >= nu) {
      const double* pm = side[0][r];  // row mu: atom A
      const double* pn = side[1][c];  // column nu: atom B
      const double s = s4[k];
      double bsum = s * (pm[0] + pn[0]);
      for (int x = 0; x < 3; x++) bsum += d4[k][x] * pn[1 + x];
      for (int x = 0; x < 6; x++) bsum += q4[k][x] * pn[4 + x];
      // moments of (nu, mu): around the atom of mu
      double vec[3] = {pn[10] - pm[10], pn[11] - pm[11], pn[12] - pm[12]};
      double dt[3], qt[6];
      shift_to_bra(vec, s, d4[k], q4[k], dt, qt);
      for (int x = 0; x < 3; x++) bsum += dt[x] * pm[1 + x];
      for (int x = 0; x < 6; x++) bsum += qt[x] * pm[4 + x];
      f = h4[k] - 0.5 * bsum;
      F[M.mat0 + mu + (long long)nu * n] = f;
    }
    tt[c][r] = f;
  }
  __syncthreads();
  // the transpose: element (32 J + r, 32 I + c) = F(mu = 32 I + c, nu = 32 J + r)
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int c = c0 + 8 * k;
    const int row = 32 * J + r, col = 32 * I + c;
    if (row < n && col < n && col >= row && (I != J || col > row)) F[M.mat0 + row + (long long)col * n] = tt[r][c];
  }
}

// Fermi filling of each molecule (base/src/occupations.rs:94-125, Newton::XTB680), entropy term
// -TS (:179-181) and the occupations focc = f_alpha + f_beta. One workgroup per molecule.
// st: per molecule {efermi, ts} into dst[2*m]; kocc (last index with focc > 1e-16, +1) into kocc.
__device__ double fermi_fill(const double* w, int n, double nel, double kt, double* occ, double* red, double& ef_out,
                             bool& ok) {
  // aufbau start (occupations.rs:55-67): midway between w[i0] and w[i1]; taken as max w[0..i0]
  // and min w[i1..n-1], which are those for sorted w and the highest occupied and lowest empty
  // level for orbitals refined for the occupied / empty split (occupied first, each block in the
  // order of the previous iteration)
  int full = (int)floor(nel);
  double rem = fmod(nel, 1.0);
  int homo = full + (rem > 0.5 ? 1 : 0);
  int i0 = max(homo, 1) - 1, i1 = min(homo + 1, n) - 1;
  double whi = -INFINITY, wlo = INFINITY;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (i <= i0) whi = fmax(whi, w[i]);
    if (i >= i1) wlo = fmin(wlo, w[i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    whi = fmax(whi, __shfl_xor(whi, o, 64));
    wlo = fmin(wlo, __shfl_xor(wlo, o, 64));
  }
  __shared__ double wsel[2][16];
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    wsel[0][threadIdx.x >> 6] = whi;
    wsel[1][threadIdx.x >> 6] = wlo;
  }
  __syncthreads();
  for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
    whi = fmax(whi, wsel[0][q]);
    wlo = fmin(wlo, wsel[1][q]);
  }
  __syncthreads();
  double ef = 0.5 * (whi + wlo), ef_used = ef;
  for (int it = 0; it < 200; it++) {
    ef_used = ef;
    double tot = 0.0, dtot = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      double arg = (w[i] - ef) / kt, f, df;
      if (arg > 50.0) {
        f = 0.0;
        df = 0.0;
      } else if (arg < -50.0) {
        f = 1.0;
        df = 0.0;
      } else {
        f = 1.0 / (exp(arg) + 1.0);
        df = -f * (1.0 - f) / kt;
      }
      occ[i] = f;
      tot += f;
      dtot -= df;
    }
    tot = block_sum(tot, red);
    dtot = block_sum(dtot, red);
    if (fabs(nel - tot) <= 1e5 * 2.220446049250313e-16) break;
    if (dtot > 1.4916681462400413e-154)
      ef += (nel - tot) / dtot;
    else
      break;
  }
  double placed = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) placed += occ[i];
  placed = block_sum(placed, red);
  ok = fabs(placed - nel) <= 1e-6;
  ef_out = ef_used;
  double ts = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double o = occ[i];
    ts += log(pow(o, o) * pow(1.0 - o, 1.0 - o));
  }
  return block_sum(ts, red) * kt;
}

__global__ void k_fermi(Batch b, const int* active, const double* w, double* focc, double* work, double* st, int* kocc,
                        int* status) {
  __shared__ double red[16];
  int m = blockIdx.x;
  if (active && !active[m]) return;
  const Mol M = b.mol[m];
  int n = M.nao;
  const double* wm = w + M.ao0;
  double* fo = focc + M.ao0;
  double* oa = work + 2 * (long long)M.ao0;
  double* ob = oa + n;
  double efa = 0.0, ts = 0.0;
  bool ok = true;
  if (M.kt > 0.0) {
    bool oka, okb;
    double efb;
    ts = fermi_fill(wm, n, M.nalpha, M.kt, oa, red, efa, oka);
    ts += fermi_fill(wm, n, M.nbeta, M.kt, ob, red, efb, okb);
    ok = oka && okb;
  } else {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      double fa = i < (int)floor(M.nalpha) ? 1.0 : (i == (int)floor(M.nalpha) ? fmod(M.nalpha, 1.0) : 0.0);
      double fb = i < (int)floor(M.nbeta) ? 1.0 : (i == (int)floor(M.nbeta) ? fmod(M.nbeta, 1.0) : 0.0);
      oa[i] = fa;
      ob[i] = fb;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) fo[i] = oa[i] + ob[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    int k = 0;
    for (int i = n - 1; i >= 0; i--)
      if (fo[i] > 1e-16) {
        k = i + 1;
        break;
      }
    kocc[m] = k;
    st[2 * m] = efa;
    st[2 * m + 1] = ts;
    if (!ok) atomicOr(&status[m], 2);
  }
}

// Occupation class of every orbital of the molecules in mlist (one workgroup each): 0 if
// f > 2 - 1e-12, 1 if f < 1e-12, else 2, at the orbital's AO index; nfrac[m] counts the class-2
// orbitals of molecule m.
__global__ void k_occ_class_batch(Batch b, const int* mlist, const double* focc, int* cls, int* nfrac) {
  __shared__ int cnt;
  const int m = mlist[blockIdx.x];
  const Mol M = b.mol[m];
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int c2 = 0;
  for (int i = threadIdx.x; i < M.nao; i += blockDim.x) {
    const double f = focc[M.ao0 + i];
    const int c = f > 2.0 - 1e-12 ? 0 : (f < 1e-12 ? 1 : 2);
    cls[M.ao0 + i] = c;
    c2 += c == 2 ? 1 : 0;
  }
  if (c2) atomicAdd(&cnt, c2);
  __syncthreads();
  if (threadIdx.x == 0) nfrac[m] = cnt;
}

// Mulliken column sums (scf/mulliken.rs:19-59) and the band energy Tr(H0 P) per column: one
// wavefront per AO column nu writes cols[11*nu + x] = sum_mu P(mu,nu) M_x(mu,nu) for
// M = S, D (3), Q (6), H0.
__global__ void k_mulliken_cols(Batch b, const int* active, const double* pool, long long stride, const double* P,
                                double* cols) {
  int g = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (g >= b.nao) return;
  int m = b.a.mol[b.ao_at[g]];
  if (active && !active[m]) return;
  const Mol M = b.mol[m];
  int nu = g - M.ao0;
  long long c0 = M.mat0 + (long long)nu * M.nao;
  double r[11];
  for (int x = 0; x < 11; x++) r[x] = 0.0;
  for (int mu = lane; mu < M.nao; mu += 64) {
    long long e = c0 + mu;
    double p = P[e];
    r[0] += p * pool[e];
    for (int k = 0; k < 3; k++) r[1 + k] += p * pool[(2 + k) * stride + e];
    for (int k = 0; k < 6; k++) r[4 + k] += p * pool[(5 + k) * stride + e];
    r[10] += p * pool[stride + e];
  }
  // the eleven wavefront sums step by step together (independent shuffles in flight)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int x = 0; x < 11; x++) r[x] += __shfl_xor(r[x], o, 64);
  if (lane == 0)
    for (int x = 0; x < 11; x++) cols[11LL * g + x] = r[x];
}

// Shell charges, atomic dipoles/quadrupoles and band energy per atom from the column sums,
// summed in AO order (one thread per atom; AOs of an atom are contiguous).
__global__ void k_mulliken_atoms(Batch b, const int* active, const double* cols, double* qsh, double* qat, double* dp,
                                 double* qp, double* eband) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= b.nat) return;
  int m = b.a.mol[a];
  if (active && !active[m]) return;
  double d[3] = {0, 0, 0}, q[6] = {0, 0, 0, 0, 0, 0}, eb = 0.0, qa = 0.0;
  const Mol M = b.mol[m];
  for (int s = b.a.ish[a]; s < b.a.ish[a] + b.a.nsh[a]; s++) {
    double pop = 0.0;
    int nao = 2 * b.s.l[s] + 1;
    for (int k = 0; k < nao; k++) {
      long long g = M.ao0 + b.s.iao[s] + k;
      const double* c = &cols[11 * g];
      pop += c[0];
      for (int x = 0; x < 3; x++) d[x] -= c[1 + x];
      for (int x = 0; x < 6; x++) q[x] -= c[4 + x];
      eb += c[10];
    }
    qsh[s] = b.s.refocc[s] - pop;
    qa += qsh[s];
  }
  qat[a] = qa;
  for (int x = 0; x < 3; x++) dp[3 * a + x] = d[x];
  for (int x = 0; x < 6; x++) qp[6 * a + x] = q[x];
  eband[a] = eb;
}

// Electronic energy per atom of the output moments (xtb.rs:1058-1146): ES2 from gq = gamma q_out,
// ES3, AES (coulomb.rs:668-731), D4 two-body (disp.rs:380-399), -TS / nat and the band energy.
// One wavefront per atom.
__global__ __launch_bounds__(256) void k_energy_atoms(Batch b, const int* active, const double* mrad, const double* gq, const double* qsh,
                               const double* qat, const double* dp, const double* qp, const double* gw,
                               const double* dgw, const double* eband, const double* st, double* eel, double* eterm) {
  int a = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (a >= b.nat) return;
  int m = b.a.mol[a];
  if (active && !active[m]) return;
  const Mol M = b.mol[m];
  const double* xa = &b.a.xyz[3 * a];
  double eaes = 0.0, ed4 = 0.0;
  for (int j = M.at0 + lane; j < M.at0 + M.nat; j += 64) {
    if (j == a) continue;
    const double* xj = &b.a.xyz[3 * j];
    double v[3] = {xa[0] - xj[0], xa[1] - xj[1], xa[2] - xj[2]};
    double r2 = dot3(v, v);
    double g3, g5;
    aes_g(sqrt(r2), mrad[a], mrad[j], g3, g5);
    const double *di = &dp[3 * a], *dj = &dp[3 * j];
    double miv = dot3(di, v), mjv = dot3(dj, v);
    double en = (qat[a] * mjv - qat[j] * miv) * g3 + g5 * (r2 * dot3(di, dj) - 3.0 * miv * mjv);
    double tc[6] = {v[0] * v[0], v[0] * v[1], v[1] * v[1], v[0] * v[2], v[1] * v[2], v[2] * v[2]};
    for (int k = 0; k < 6; k++) en += QSCALE[k] * tc[k] * g5 * (qat[j] * qp[6 * a + k] + qat[a] * qp[6 * j + k]);
    eaes += 0.5 * en;
    double pe, pg;
    if (d4_pair_e(b, a, j, v, r2, pe, pg)) {
      double c6, x, y;
      pair_c6(b, a, j, gw, dgw, c6, x, y);
      ed4 += -(c6 * pe * 0.5);
    }
  }
  eaes = wave_sum(eaes);
  ed4 = wave_sum(ed4);
  if (lane == 0) {
    double es2 = 0.0, es3 = 0.0;
    for (int s = b.a.ish[a]; s < b.a.ish[a] + b.a.nsh[a]; s++) {
      es2 += 0.5 * qsh[s] * gq[s];
      es3 += qsh[s] * qsh[s] * qsh[s] * b.s.gam3[s] / 3.0;
    }
    const double* d = &dp[3 * a];
    eaes += b.a.dkernel[a] * dot3(d, d);
    double qq = 0.0;
    for (int k = 0; k < 6; k++) qq += QSCALE[k] * qp[6 * a + k] * qp[6 * a + k];
    eaes += b.a.qkernel[a] * qq;
    double ts = st[2 * m + 1];
    eel[a] = es2 + es3 + eaes + ed4 + ts / M.nat + eband[a];
    double* t = &eterm[8 * (long long)a];
    t[0] = es2;
    t[1] = es3;
    t[2] = eaes;
    t[3] = ed4;
    t[4] = eband[a];
  }
}

// Per molecule: the energy (sum of atoms in order), the residual norm of dq, the convergence
// test (scf/convergence.rs:99-121) and the iteration record. One workgroup per molecule.
// mst[m*8]: {energy, energy_change, residual, efermi, ts, iter, converged, -}
__global__ void k_converge(Batch b, int* active, const double* eel, const double* dq, const double* st, double econv,
                           double pconv, double* mst, int* status) {
  __shared__ double red[16];
  int m = blockIdx.x;
  if (!active[m]) return;
  const Mol M = b.mol[m];
  double e = 0.0;
  if (threadIdx.x == 0)
    for (int a = M.at0; a < M.at0 + M.nat; a++) e += eel[a];
  double r = 0.0;
  for (int k = threadIdx.x; k < M.ndim; k += blockDim.x) {
    double x = dq[M.vec0 + k];
    r += x * x;
  }
  r = block_sum(r, red);
  if (threadIdx.x == 0) {
    double* s = &mst[8 * m];
    double res = sqrt(r / M.ndim);
    double de = e - s[0];
    s[0] = e;
    s[1] = de;
    s[2] = res;
    s[3] = st[2 * m];
    s[4] = st[2 * m + 1];
    s[5] += 1.0;
    if (!isfinite(e)) {
      atomicOr(&status[m], 4);
      active[m] = 0;
    } else if (fabs(de) < econv && res < pconv) {
      s[6] = 1.0;
      active[m] = 0;
    } else if (status[m]) {
      active[m] = 0;
    }
  }
}

// ------------------------------------------------------------------ Broyden mixing
// Modified Broyden mixing of one molecule per workgroup (scf/mixing.rs:251-338, omega0 0.01,
// minw 1, maxw 1e5, wfac 0.01) with the reference's full history (`memory` = max_iter slots,
// reused cyclically beyond it). The history is laid out slot-major so that it grows with the
// iterations without moving: df, u at slot * tot_vec + vec0, omega at slot * nmol + m, the
// symmetric overlaps a(i, j) packed as (max (max + 1) / 2 + min) * nmol + m. The small system
// (ld x ld, ld >= the history length) is solved by Gaussian elimination with the last-maximum
// pivot (linalg/lu.rs:81-107), rows eliminated in parallel; it lives in LDS, or in `gbeta`
// (ld * ld per molecule) when it does not fit.
__device__ __forceinline__ long long tri_index(int i, int j) {
  const int hi = i > j ? i : j, lo = i > j ? j : i;
  return (long long)hi * (hi + 1) / 2 + lo;
}

__global__ void k_broyden(Batch b, const int* active, int cap, int ld, double damp, double* qin, const double* dq,
                          double* qlast, double* dqlast, double* df, double* u, long long tot_vec, double* omega,
                          double* amat, int nmol, double* gbeta, int* itc, int* status) {
  extern __shared__ double sm[];
  int m = blockIdx.x;
  if (!active[m]) return;
  const Mol M = b.mol[m];
  const int n = M.ndim;
  double* q = qin + M.vec0;
  const double* d = dq + M.vec0;
  double* ql = qlast + M.vec0;
  double* dl = dqlast + M.vec0;
  double* red = sm;                       // 16
  double* cvec = sm + 16;                 // ld
  int* hist = (int*)(cvec + ld);          // ld ints (ld doubles reserved)
  double* beta = gbeta ? gbeta + (long long)m * ld * ld : cvec + 2 * ld;
  __shared__ int it_sh, ok_sh, piv_sh;
  if (threadIdx.x == 0) it_sh = ++itc[m];
  __syncthreads();
  const int iter = it_sh;
  if (iter == 1) {
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      dl[k] = d[k];
      ql[k] = q[k];
      q[k] += damp * d[k];
    }
    return;
  }
  const int itn = iter - 1;
  const int it1 = (itn - 1) % cap;
  const int nhist = min(cap, itn);
  if (nhist > ld) {  // the host sizes ld for the iteration; never taken
    if (threadIdx.x == 0) atomicOr(&status[m], 8);
    return;
  }
  const int first = itn + 1 > cap ? itn + 1 - cap : 1;
  const int th = threadIdx.x, bd = blockDim.x;
  for (int ih = th; ih < nhist; ih += bd) hist[ih] = (first + ih - 1) % cap;
  double* dfi = df + (long long)it1 * tot_vec + M.vec0;
  double* ui = u + (long long)it1 * tot_vec + M.vec0;
  double nn = 0.0;
  for (int k = th; k < n; k += bd) nn += d[k] * d[k];
  nn = block_sum(nn, red);
  double norm = sqrt(nn);
  double om = norm > 0.01 / 100000.0 ? 0.01 / norm : 100000.0;
  if (om < 1.0) om = 1.0;
  double ff = 0.0;
  for (int k = th; k < n; k += bd) {
    double x = d[k] - dl[k];
    dfi[k] = x;
    ff += x * x;
  }
  ff = block_sum(ff, red);
  double inv = 1.0 / fmax(sqrt(ff), 2.220446049250313e-16);
  for (int k = th; k < n; k += bd) dfi[k] *= inv;
  if (th == 0) omega[(long long)it1 * nmol + m] = om;
  __syncthreads();
  // a(i, it1) for the history slots and c
  for (int ih = 0; ih < nhist; ih++) {
    int i = hist[ih];
    const double* dfh = df + (long long)i * tot_vec + M.vec0;
    double s1 = 0.0, s2 = 0.0;
    for (int k = th; k < n; k += bd) {
      s1 += dfh[k] * d[k];
      s2 += dfh[k] * dfi[k];
    }
    s1 = block_sum(s1, red);
    s2 = block_sum(s2, red);
    if (th == 0) {
      cvec[ih] = omega[(long long)i * nmol + m] * s1;
      amat[tri_index(i, it1) * nmol + m] = s2;
    }
  }
  __syncthreads();
  for (int x = th; x < nhist * nhist; x += bd) {
    const int ih = x / nhist, jh = x - ih * nhist;
    const int i = hist[ih], j = hist[jh];
    double v = omega[(long long)i * nmol + m] * omega[(long long)j * nmol + m] * amat[tri_index(i, j) * nmol + m];
    if (ih == jh) v += 0.01 * 0.01;
    beta[ih * ld + jh] = v;
  }
  if (th == 0) ok_sh = 1;
  __syncthreads();
  // Gaussian elimination, pivot on the last maximum of |beta(i, col)|; each row's operations are
  // those of the sequential elimination
  for (int col = 0; col < nhist; col++) {
    if (th == 0) {
      int piv = col;
      for (int i = col; i < nhist; i++)
        if (fabs(beta[i * ld + col]) >= fabs(beta[piv * ld + col])) piv = i;
      piv_sh = piv;
      if (beta[piv * ld + col] == 0.0) ok_sh = 0;
    }
    __syncthreads();
    if (!ok_sh) break;
    const int piv = piv_sh;
    if (piv != col) {
      for (int k = th; k < nhist; k += bd) {
        double t = beta[col * ld + k];
        beta[col * ld + k] = beta[piv * ld + k];
        beta[piv * ld + k] = t;
      }
      if (th == 0) {
        double t = cvec[col];
        cvec[col] = cvec[piv];
        cvec[piv] = t;
      }
    }
    __syncthreads();
    for (int row = col + 1 + th; row < nhist; row += bd) {
      double f = beta[row * ld + col] / beta[col * ld + col];
      if (f != 0.0) {
        for (int k = col; k < nhist; k++) beta[row * ld + k] -= f * beta[col * ld + k];
        cvec[row] -= f * cvec[col];
      }
    }
    __syncthreads();
  }
  if (th == 0) {
    if (ok_sh)
      for (int col = nhist - 1; col >= 0; col--) {
        double s = cvec[col];
        for (int k = col + 1; k < nhist; k++) s -= beta[col * ld + k] * cvec[k];
        cvec[col] = s / beta[col * ld + col];
      }
    else
      atomicOr(&status[m], 8);
  }
  __syncthreads();
  if (!ok_sh) return;
  for (int k = th; k < n; k += bd) {
    double uk = damp * dfi[k] + inv * (q[k] - ql[k]);
    ui[k] = uk;
    dl[k] = d[k];
    ql[k] = q[k];
    double x = q[k] + damp * d[k];
    for (int ih = 0; ih < nhist; ih++) {
      int i = hist[ih];
      double uv = (i == it1) ? uk : u[(long long)i * tot_vec + M.vec0 + k];
      x -= omega[(long long)i * nmol + m] * cvec[ih] * uv;
    }
    q[k] = x;
  }
}

// ------------------------------------------------------------------ DIIS mixing
// Pulay (DIIS) mixing with restarts, one molecule per workgroup (scf/mixing.rs:340-513): the last
// DIIS_SIZE inputs and residuals in a ring (slot-major, slot * tot_vec + vec0); the coefficients
// minimise |sum c_i r_i| with sum c_i = 1, from the differences to the newest residual
// orthogonalised newest first by modified Gram-Schmidt (a difference whose new part is below
// DIIS_DEPENDENCE of its norm ends the basis); a residual above DIIS_RESTART_GROWTH times the best
// one (and above DIIS_RESTART_FLOOR), or DIIS_STALL iterations without a better one, restarts from
// the best input with half the damping. State per molecule: dst[2 m] damping, dst[2 m + 1] best
// residual norm; ist[4 m ..] has-best, iterations since the best, history length, ring head.
constexpr int DIIS_SIZE = 30;
constexpr double DIIS_DEPENDENCE = 1e-2;
constexpr int DIIS_STALL = 12;
constexpr double DIIS_RESTART_GROWTH = 1e3;
constexpr double DIIS_RESTART_FLOOR = 0.1;

__global__ void k_diis(Batch b, const int* active, double* qin, const double* dq, double* hq, double* hr, double* bq,
                       double* br, double* bas, long long tot_vec, double* dst, int* ist, int* status) {
  __shared__ double red[16];
  __shared__ double rcol[DIIS_SIZE][DIIS_SIZE];  // rcol[p][l]: column p of R
  __shared__ double rhs[DIIS_SIZE], cf[DIIS_SIZE];
  __shared__ int kept[DIIS_SIZE];
  const int m = blockIdx.x;
  if (!active[m]) return;
  const Mol M = b.mol[m];
  const int n = M.ndim, th = threadIdx.x, bd = blockDim.x;
  double* q = qin + M.vec0;
  const double* d = dq + M.vec0;
  auto vslot = [&](double* base, int s) { return base + (long long)s * tot_vec + M.vec0; };
  double nn = 0.0;
  for (int k = th; k < n; k += bd) nn += d[k] * d[k];
  const double norm = sqrt(block_sum(nn, red));
  if (!isfinite(norm)) {
    if (th == 0) atomicOr(&status[m], 8);
    return;
  }
  double damp = dst[2 * m], best = dst[2 * m + 1];
  int has = ist[4 * m], since = ist[4 * m + 1], cnt = ist[4 * m + 2], head = ist[4 * m + 3];
  double* qb = bq + M.vec0;
  double* rb = br + M.vec0;
  if (has) {
    const bool diverging = norm > DIIS_RESTART_GROWTH * best && norm > DIIS_RESTART_FLOOR;
    if (diverging || since >= DIIS_STALL) {
      damp *= 0.5;
      for (int k = th; k < n; k += bd) q[k] = qb[k] + damp * rb[k];
      if (th == 0) {
        dst[2 * m] = damp;
        ist[4 * m + 1] = 0;
        ist[4 * m + 2] = 0;
      }
      return;
    }
  }
  if (!has || norm < best) {
    best = norm;
    has = 1;
    since = 0;
    for (int k = th; k < n; k += bd) {
      qb[k] = q[k];
      rb[k] = d[k];
    }
  } else {
    since++;
  }
  if (cnt == DIIS_SIZE) {
    head = (head + 1) % DIIS_SIZE;
    cnt--;
  }
  {
    const int s = (head + cnt) % DIIS_SIZE;
    double* hqs = vslot(hq, s);
    double* hrs = vslot(hr, s);
    for (int k = th; k < n; k += bd) {
      hqs[k] = q[k];
      hrs[k] = d[k];
    }
    cnt++;
  }
  // every thread owns the vector entries k = th + j bd: the element-wise passes need no barrier
  const double* last = vslot(hr, (head + cnt - 1) % DIIS_SIZE);
  int p = 0;
  for (int jj = cnt - 2; jj >= 0; jj--) {
    const double* rj = vslot(hr, (head + jj) % DIIS_SIZE);
    double* dv = vslot(bas, p);
    double s = 0.0;
    for (int k = th; k < n; k += bd) {
      const double x = last[k] - rj[k];
      dv[k] = x;
      s += x * x;
    }
    const double norm0 = sqrt(block_sum(s, red));
    if (isnan(norm0) || norm0 <= 0.0) break;
    for (int l = 0; l < p; l++) {
      const double* bl = vslot(bas, l);
      double h = 0.0;
      for (int k = th; k < n; k += bd) h += bl[k] * dv[k];
      h = block_sum(h, red);
      if (th == 0) rcol[p][l] = h;
      for (int k = th; k < n; k += bd) dv[k] -= h * bl[k];
    }
    double r2 = 0.0;
    for (int k = th; k < n; k += bd) r2 += dv[k] * dv[k];
    const double nr = sqrt(block_sum(r2, red));
    if (nr < DIIS_DEPENDENCE * norm0) break;
    if (th == 0) {
      rcol[p][p] = nr;
      kept[p] = jj;
    }
    for (int k = th; k < n; k += bd) dv[k] /= nr;
    p++;
  }
  for (int l = 0; l < p; l++) {
    const double* bl = vslot(bas, l);
    double h = 0.0;
    for (int k = th; k < n; k += bd) h += bl[k] * last[k];
    h = block_sum(h, red);
    if (th == 0) rhs[l] = h;
  }
  __syncthreads();
  if (th == 0) {
    double g[DIIS_SIZE];
    for (int l = p - 1; l >= 0; l--) {
      double v = rhs[l];
      for (int k = l + 1; k < p; k++) v -= rcol[k][l] * g[k];
      g[l] = v / rcol[l][l];
    }
    for (int i = 0; i < cnt; i++) cf[i] = 0.0;
    cf[cnt - 1] = 1.0;
    for (int l = 0; l < p; l++) {
      cf[kept[l]] += g[l];
      cf[cnt - 1] -= g[l];
    }
  }
  __syncthreads();
  for (int k = th; k < n; k += bd) {
    double x = 0.0;
    for (int i = 0; i < cnt; i++) {
      const double c = cf[i];
      if (c != 0.0) {
        const int s = (head + i) % DIIS_SIZE;
        x += c * (vslot(hq, s)[k] + damp * vslot(hr, s)[k]);
      }
    }
    q[k] = x;
  }
  if (th == 0) {
    dst[2 * m] = damp;
    dst[2 * m + 1] = best;
    ist[4 * m] = has;
    ist[4 * m + 1] = since;
    ist[4 * m + 2] = cnt;
    ist[4 * m + 3] = head;
  }
}

// The start of an SCF attempt for the molecules of `mask` (their iteration record, status and
// mixer state cleared; DIIS at the given damping).
__global__ void k_attempt_start(const int* mask, int nmol, int* active, int* itc, int* status, double* mst,
                                double* dst, int* ist, double damp) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol) return;
  active[m] = mask[m];
  if (!mask[m]) return;
  itc[m] = 0;
  status[m] = 0;
  for (int k = 0; k < 8; k++) mst[8 * m + k] = 0.0;
  if (dst) {
    dst[2 * m] = damp;
    dst[2 * m + 1] = 0.0;
    for (int k = 0; k < 4; k++) ist[4 * m + k] = 0;
  }
}

// A Jacobi solve that did not converge within its sweeps (stat < 0) marks its molecule with
// status bit 16 (Error::Numerical of the eigensolver).
__global__ void k_jacobi_flag(const int* stat, const int* mol_of_job, int njobs, int* status) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < njobs && stat[q] < 0 && stat[q] != JACOBI_SKIPPED) atomicOr(&status[mol_of_job[q]], 16);
}

// ------------------------------------------------------------------ gradient
// The classical pair gradients of atom a (one wavefront per atom; every pair is evaluated from
// both atoms, each keeping its own share): ES2 (coulomb.rs:123-169), AES with dE/dCN through the
// damping radii (coulomb.rs:896-1001) and D4 two-body at fixed charges with dE/dCN_D4
// (disp.rs:460-484); plus the atom's share of the integral-derivative pair terms of k_h0_grad and
// the on-site dE/dCN of its levels (hamiltonian.rs:212-221).
__global__ __launch_bounds__(256) void k_pair_grad(Batch b, const int* okm, const double* pool, long long stride, const double* P,
                            const double* W, const double* se, const double* dsedcn, const double* vsh,
                            const double* vat, const double* vdp, const double* vqp, const double* qsh,
                            const double* qat, const double* dp, const double* qp, const double* mrad,
                            const double* dmrdcn, const double* gw, const double* dgwdcn, const double* dcn3,
                            const double* hg, long long hstride, double* g, double* dedcn, double* dedcn_d4) {
  int a = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (a >= b.nat) return;
  int m = b.a.mol[a];
  if (okm && !okm[m]) return;
  const Mol M = b.mol[m];
  double ga[3] = {0, 0, 0}, dcn = 0.0, dcn4 = 0.0, dedr = 0.0;
  const double kdmp3 = 3.0, kdmp5 = 4.0;
  for (int j = M.at0 + lane; j < M.at0 + M.nat; j += 64) {
    if (j == a) continue;
    const int iat = a > j ? a : j, jat = a > j ? j : a;
    const double sgn = a == iat ? 1.0 : -1.0;
    const double* xi = &b.a.xyz[3 * iat];
    const double* xj = &b.a.xyz[3 * jat];
    double v[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};  // R_iat - R_jat
    double r2 = dot3(v, v), r1 = sqrt(r2);
    // ES2
    {
      double dgs = 0.0;
      for (int s = b.a.ish[iat]; s < b.a.ish[iat] + b.a.nsh[iat]; s++)
        for (int t = b.a.ish[jat]; t < b.a.ish[jat] + b.a.nsh[jat]; t++) {
          double eta = 0.5 * (b.s.hard[s] + b.s.hard[t]);
          double dt = 1.0 / (r2 + 1.0 / (eta * eta));
          dgs += -dt * sqrt(dt) * qsh[s] * qsh[t];
        }
      for (int k = 0; k < 3; k++) ga[k] += sgn * dgs * v[k];
    }
    // AES (vec = R_jat - R_iat)
    {
      double vec[3] = {-v[0], -v[1], -v[2]};
      double rr = 0.5 * (mrad[iat] + mrad[jat]);
      double g1 = 1.0 / r1, g3 = g1 * g1 * g1, g5 = g3 * g1 * g1, g7 = g5 * g1 * g1;
      double x = rr * g1, x3 = x * x * x, x4 = x3 * x;
      double fdmp3 = 1.0 / (1.0 + 6.0 * x3);
      double ddmp3 = -3.0 * g5 * fdmp3 - kdmp3 * fdmp3 * (fdmp3 - 1.0) * g5;
      double fdmp5 = 1.0 / (1.0 + 6.0 * x4);
      double ddmp5 = -5.0 * fdmp5 - kdmp5 * (fdmp5 * fdmp5 - fdmp5);
      double qi = qat[iat], qj = qat[jat];
      const double *di = &dp[3 * iat], *dj = &dp[3 * jat];
      double dpiqj = dot3(vec, di) * qj, qidpj = dot3(vec, dj) * qi;
      double fddr = 3.0 * (dpiqj - qidpj) * kdmp3 * fdmp3 * g3 * (fdmp3 / rr) * x3;
      double dg[3];
      for (int k = 0; k < 3; k++) dg[k] = -ddmp3 * vec[k] * (dpiqj - qidpj) + fdmp3 * g3 * (qi * dj[k] - qj * di[k]);
      double dpidpj = dot3(dj, di), dpiv = dot3(di, vec), dpjv = dot3(dj, vec);
      double edd = dpidpj * r2 - 3.0 * dpjv * dpiv;
      fddr += 3.0 * edd * kdmp5 * fdmp5 * g5 * (fdmp5 / rr) * x4;
      for (int k = 0; k < 3; k++)
        dg[k] += -2.0 * fdmp5 * g5 * dpidpj * vec[k] + 3.0 * fdmp5 * g5 * (dpiv * dj[k] + dpjv * di[k]) -
                 edd * ddmp5 * g7 * vec[k];
      const double *ti = &qp[6 * iat], *tj = &qp[6 * jat];
      double eq = 2.0 * (qj * ti[1] + tj[1] * qi) * vec[0] * vec[1] + 2.0 * (qj * ti[3] + tj[3] * qi) * vec[0] * vec[2] +
                  2.0 * (qj * ti[4] + tj[4] * qi) * vec[1] * vec[2] + (qj * ti[0] + tj[0] * qi) * vec[0] * vec[0] +
                  (qj * ti[2] + tj[2] * qi) * vec[1] * vec[1] + (qj * ti[5] + tj[5] * qi) * vec[2] * vec[2];
      fddr += eq * 3.0 * kdmp5 * fdmp5 * g5 * fdmp5 / rr * x4;
      double tvj[3] = {vec[0] * tj[0] + vec[1] * tj[1] + vec[2] * tj[3], vec[0] * tj[1] + vec[1] * tj[2] + vec[2] * tj[4],
                       vec[0] * tj[3] + vec[1] * tj[4] + vec[2] * tj[5]};
      double tvi[3] = {vec[0] * ti[0] + vec[1] * ti[1] + vec[2] * ti[3], vec[0] * ti[1] + vec[1] * ti[2] + vec[2] * ti[4],
                       vec[0] * ti[3] + vec[1] * ti[4] + vec[2] * ti[5]};
      for (int k = 0; k < 3; k++)
        dg[k] += -eq * ddmp5 * g7 * vec[k] - 2.0 * fdmp5 * g5 * qi * tvj[k] - 2.0 * fdmp5 * g5 * qj * tvi[k];
      for (int k = 0; k < 3; k++) ga[k] += sgn * dg[k];
      dedr += fddr;
    }
    // D4 two-body at fixed charges
    {
      double pe, pg;
      if (d4_pair_e(b, iat, jat, v, r2, pe, pg)) {
        double c6, dci, dcj;
        pair_c6(b, iat, jat, gw, dgwdcn, c6, dci, dcj);
        dcn4 -= (a == iat ? dci : dcj) * pe;
        for (int k = 0; k < 3; k++) ga[k] += sgn * (-c6 * pg * v[k]);
      }
    }
  }
  // the integral-derivative terms of the pairs (a, j < a) and (i > a, a) from k_h0_grad
  {
    const int la = a - M.at0;
    for (int lj = lane; lj < M.nat; lj += 64) {
      if (lj == la) continue;
      long long e = la > lj ? M.pair0 + (long long)la * M.nat + lj : M.pair0 + (long long)lj * M.nat + la;
      double sg = la > lj ? 1.0 : -1.0;
      for (int k = 0; k < 3; k++) ga[k] += sg * hg[k * hstride + e];
      dcn += la > lj ? hg[3 * hstride + e] : hg[4 * hstride + e];
    }
  }
  for (int k = 0; k < 3; k++) ga[k] = wave_sum(ga[k]);
  dcn = wave_sum(dcn);
  dcn4 = wave_sum(dcn4);
  dedr = wave_sum(dedr);
  if (lane == 0) {
    // on-site: the CN dependence of the levels (hamiltonian.rs:212-221)
    for (int s = b.a.ish[a]; s < b.a.ish[a] + b.a.nsh[a]; s++) {
      int n = 2 * b.s.l[s] + 1;
      for (int k = 0; k < n; k++) {
        long long mu = b.s.iao[s] + k;
        dcn += dsedcn[s] * P[M.mat0 + mu + mu * M.nao];
      }
    }
    dcn += dedr * dmrdcn[a];
    for (int k = 0; k < 3; k++) g[3 * a + k] += ga[k];
    dedcn[a] = dcn;
    dedcn_d4[a] = dcn4 + dcn3[a];
  }
}

// H0 / S / multipole integral derivatives of one atom pair (iat > jat) per thread, over its
// screened shell pairs, contracted with P, W, the AO potentials and the multipole potentials
// (hamiltonian.rs:137-223). Writes hg[x * hstride + pair0 + iat*nat + jat] for x = dg (3),
// dE/dCN of iat, dE/dCN of jat.
// Block coefficients of the ket rows [IM0, IM0 + NR) of a shell pair (hamiltonian.rs:169-211) and
// their contraction with the integral derivatives; every row chunk of the type in turn (all rows
// at once for s and p pairs). Returns sum_blk p s.
template <int LI, int LJ, int IM0, int NR>
__device__ __forceinline__ double h0_grad_block(const Batch& b, const Mol& M, int si, int sj, int iat, int jat,
                                                const double* v, double hij, const double* dsv, const double* P,
                                                const double* W, const double* vsh, const double* vat,
                                                const double* vdp, const double* vqp, double* dg) {
  constexpr int NSJ = 2 * LJ + 1;
  double sval[NR * NSJ], hp2[NR * NSJ], pp[NR * NSJ];
  const int ii = b.s.iao[si], jj = b.s.iao[sj];
  const double vi = vsh[si] + vat[iat], vj = vsh[sj] + vat[jat];
#pragma unroll
  for (int r = 0; r < NR; r++)
#pragma unroll
    for (int jm = 0; jm < NSJ; jm++) {
      const int blk = jm + NSJ * r;
      long long e = M.mat0 + (jj + jm) + (long long)(ii + IM0 + r) * M.nao;
      double pij = P[e];
      pp[blk] = pij;
      hp2[blk] = 2.0 * pij * hij;
      sval[blk] = -pij * (vj + vi) + 2.0 * pij * hij - 2.0 * W[e];
    }
  return shellpair_grad_t<LI, LJ, IM0, NR>(b.s.nprim[si], &b.s.alpha[MAXPRIM * si], &b.s.coeff[MAXPRIM * si],
                                           b.s.nprim[sj], &b.s.alpha[MAXPRIM * sj], &b.s.coeff[MAXPRIM * sj], v,
                                           b.intcut, sval, hp2, pp, dsv, &vdp[3 * iat], &vdp[3 * jat], &vqp[6 * iat],
                                           &vqp[6 * jat], dg);
}

template <int LI, int LJ, int IM = 0>
__device__ __forceinline__ double h0_grad_rows(const Batch& b, const Mol& M, int si, int sj, int iat, int jat,
                                               const double* v, double hij, const double* dsv, const double* P,
                                               const double* W, const double* vsh, const double* vat,
                                               const double* vdp, const double* vqp, double* dg) {
  using RS = RowSplit<LI, LJ>;
  if constexpr (!RS::split) {
    return h0_grad_block<LI, LJ, 0, RS::NR>(b, M, si, sj, iat, jat, v, hij, dsv, P, W, vsh, vat, vdp, vqp, dg);
  } else if constexpr (IM < RS::NCH) {
    const double ps = h0_grad_block<LI, LJ, IM, 1>(b, M, si, sj, iat, jat, v, hij, dsv, P, W, vsh, vat, vdp, vqp, dg);
    return ps + h0_grad_rows<LI, LJ, IM + 1>(b, M, si, sj, iat, jat, v, hij, dsv, P, W, vsh, vat, vdp, vqp, dg);
  } else {
    return 0.0;
  }
}

template <int LI, int LJ>
__device__ __forceinline__ void h0_grad_shell(const Batch& b, const Mol& M, int si, int sj, int iat, int jat,
                                              const double* v, double r2, double rr, const double* P,
                                              const double* W, const double* se, const double* dsedcn,
                                              const double* vsh, const double* vat, const double* vdp,
                                              const double* vqp, double* dg, double& dcni, double& dcnj) {
  double shpi = 1.0 + b.s.shpoly[si] * rr, shpj = 1.0 + b.s.shpoly[sj] * rr;
  double shp = shpi * shpj;
  double dshp = (shpi * b.s.shpoly[sj] + shpj * b.s.shpoly[si]) * 0.5 * rr / r2;
  double dsv[3] = {dshp / shp * v[0], dshp / shp * v[1], dshp / shp * v[2]};
  double hs = hscale(b, sj, si) * shp;
  double hij = 0.5 * (se[si] + se[sj]) * hs;
  const double ps = h0_grad_rows<LI, LJ>(b, M, si, sj, iat, jat, v, hij, dsv, P, W, vsh, vat, vdp, vqp, dg);
  dcni += dsedcn[si] * hs * ps;
  dcnj += dsedcn[sj] * hs * ps;
}

// The H0 / overlap part of the gradient per atom pair li > lj into hg (stride hstride: dE/dR_ij (3),
// dE/dCN_i, dE/dCN_j), split by shell-pair type so that a wavefront runs one type: k_h0_keys lists,
// per key 3 l_i + l_j, the atom pairs that hold such a shell pair after the
// integral screening, in H0_GROUPS lists per key (molecule m in group m % H0_GROUPS, whose list
// starts at goff[group] within the key's region of cap entries; slots taken per wavefront, counted
// in cnt[key * H0_GROUPS + group]: many counters, little contention); k_h0_grad_key<LI, LJ>
// (grid.y = group) then adds that shell pair's terms to hg (zeroed before the first key).
// Launched in key order, every atom pair accumulates its shell pairs type by type (the same order
// on every run), within a type in the order of a loop over the shells of i, then of j.
constexpr int H0_GROUPS = 64;
__global__ __launch_bounds__(64) void k_h0_keys(Batch b, const int* okm, int* cnt, int2* list, long long cap,
                                                const long long* goff) {
  const int m = blockIdx.y;
  const Mol M = b.mol[m];
  const long long idx = (long long)blockIdx.x * 64 + threadIdx.x;
  int mask = 0;
  if ((!okm || okm[m]) && idx < (long long)M.nat * M.nat) {
    const int li = (int)(idx / M.nat), lj = (int)(idx % M.nat);
    if (li > lj) {
      const int iat = M.at0 + li, jat = M.at0 + lj;
      const double* xi = &b.a.xyz[3 * iat];
      const double* xj = &b.a.xyz[3 * jat];
      const double v[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
      const double r2 = dot3(v, v);
      double amin_i = 1e300, amin_j = 1e300;
      for (int s = b.a.ish[iat]; s < b.a.ish[iat] + b.a.nsh[iat]; s++) amin_i = fmin(amin_i, b.s.minalpha[s]);
      for (int s = b.a.ish[jat]; s < b.a.ish[jat] + b.a.nsh[jat]; s++) amin_j = fmin(amin_j, b.s.minalpha[s]);
      if (!(amin_i * amin_j * r2 / (amin_i + amin_j) > b.intcut * (1.0 + 1e-10)) && r2 >= 2.220446049250313e-16 &&
          !(r2 > M.pair_cut2))  // the conventions' atom-pair cutoff (ao.rs:96-103)
        for (int si = b.a.ish[iat]; si < b.a.ish[iat] + b.a.nsh[iat]; si++)
          for (int sj = b.a.ish[jat]; sj < b.a.ish[jat] + b.a.nsh[jat]; sj++) {
            const double a1 = b.s.minalpha[si], a2 = b.s.minalpha[sj];
            if (a1 * a2 * r2 / (a1 + a2) > b.intcut * (1.0 + 1e-10)) continue;
            mask |= 1 << (3 * b.s.l[si] + b.s.l[sj]);
          }
    }
  }
  const int lane = threadIdx.x;
  for (int k = 0; k < NKEYS; k++) {
    const bool has = (mask >> k) & 1;
    const unsigned long long bal = __ballot(has);
    if (!bal) continue;
    int base = 0;
    const int g = m % H0_GROUPS;
    if (lane == 0) base = atomicAdd(&cnt[k * H0_GROUPS + g], (int)__popcll(bal));
    base = __shfl(base, 0, 64);
    if (has) list[k * cap + goff[g] + base + __popcll(bal & ((1ull << lane) - 1ull))] = make_int2(m, (int)idx);
  }
}

template <int LI, int LJ>
__global__ __launch_bounds__(64) void k_h0_grad_key(Batch b, const int2* list, const int* count,
                                                    const long long* goff, const double* P, const double* W,
                                                    const double* se, const double* dsedcn, const double* vsh,
                                                    const double* vat, const double* vdp, const double* vqp,
                                                    double* hg, long long hstride) {
  const int g = blockIdx.y, t = blockIdx.x * 64 + threadIdx.x;
  if (t >= count[g]) return;
  const int2 it = list[goff[g] + t];
  const int m = it.x;
  const Mol M = b.mol[m];
  const int idx = it.y, li = idx / M.nat, lj = idx % M.nat;
  const int iat = M.at0 + li, jat = M.at0 + lj;
  const double* xi = &b.a.xyz[3 * iat];
  const double* xj = &b.a.xyz[3 * jat];
  double v[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
  const double r2 = dot3(v, v);
  const double rr = sqrt(sqrt(r2) / (b.a.rad[jat] + b.a.rad[iat]));
  const long long e = M.pair0 + idx;
  double dg[3] = {hg[e], hg[hstride + e], hg[2 * hstride + e]};
  double dcni = hg[3 * hstride + e], dcnj = hg[4 * hstride + e];
  for (int si = b.a.ish[iat]; si < b.a.ish[iat] + b.a.nsh[iat]; si++) {
    if (b.s.l[si] != LI) continue;
    for (int sj = b.a.ish[jat]; sj < b.a.ish[jat] + b.a.nsh[jat]; sj++) {
      if (b.s.l[sj] != LJ) continue;
      const double a1 = b.s.minalpha[si], a2 = b.s.minalpha[sj];
      if (a1 * a2 * r2 / (a1 + a2) > b.intcut * (1.0 + 1e-10)) continue;
      h0_grad_shell<LI, LJ>(b, M, si, sj, iat, jat, v, r2, rr, P, W, se, dsedcn, vsh, vat, vdp, vqp, dg, dcni, dcnj);
    }
  }
  for (int k = 0; k < 3; k++) hg[k * hstride + e] = dg[k];
  hg[3 * hstride + e] = dcni;
  hg[4 * hstride + e] = dcnj;
}
#define AXTB_H0_KEY(LI, LJ)                                                                                       \
  template __global__ void k_h0_grad_key<LI, LJ>(Batch, const int2*, const int*, const long long*, const double*, \
                                                 const double*, const double*, const double*, const double*,      \
                                                 const double*, const double*, const double*, double*, long long);
AXTB_H0_KEY(0, 0)
AXTB_H0_KEY(0, 1)
AXTB_H0_KEY(0, 2)
AXTB_H0_KEY(1, 0)
AXTB_H0_KEY(1, 1)
AXTB_H0_KEY(1, 2)
AXTB_H0_KEY(2, 0)
AXTB_H0_KEY(2, 1)
AXTB_H0_KEY(2, 2)
#undef AXTB_H0_KEY

// Chain rules of the two coordination numbers (base/src/ncoord.rs:486-519): the GFN2 CN with
// dE/dCN from the levels and the AES radii, and the D4 CN (Jacobian order, zero weights skipped)
// with dE/dCN of the two- and three-body dispersion. One wavefront per atom.
__global__ void k_cn_chain(Batch b, const int* okm, const double* dedcn, const double* dedcn_d4, double* g) {
  int a = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (a >= b.nat) return;
  int m = b.a.mol[a];
  if (okm && !okm[m]) return;
  const Mol M = b.mol[m];
  const double* xa = &b.a.xyz[3 * a];
  double ga[3] = {0, 0, 0};
  for (int j = M.at0 + lane; j < M.at0 + M.nat; j += 64) {
    if (j == a) continue;
    const double* xj = &b.a.xyz[3 * j];
    double v[3] = {xa[0] - xj[0], xa[1] - xj[1], xa[2] - xj[2]};
    double r2 = dot3(v, v);
    if (r2 < 1e-12 || (r2 > b.cn_cut2 && r2 > b.d4cn_cut2)) continue;
    double r = sqrt(r2);
    double t = r2 <= b.cn_cut2 ? dexp_dcount(r, b.a.rcov_cn[a] + b.a.rcov_cn[j]) * (dedcn[a] + dedcn[j]) / r : 0.0;
    double s4 = dedcn_d4[a] + dedcn_d4[j];
    double t4 = s4 != 0.0 && r2 <= b.d4cn_cut2
                    ? d4_en_factor(b.a.en[a], b.a.en[j]) * erf_dcount(r, b.a.d4_rcov[a] + b.a.d4_rcov[j]) * s4 / r
                    : 0.0;
    for (int k = 0; k < 3; k++) ga[k] += (t + t4) * v[k];
  }
  for (int k = 0; k < 3; k++) ga[k] = wave_sum(ga[k]);
  if (lane == 0)
    for (int k = 0; k < 3; k++) g[3 * a + k] += ga[k];
}

}  // namespace axtb
