// Shared definitions of the GFN2-xTB HIP engine: batch descriptors (structure-of-arrays over
// all atoms, shells and AOs of a batch of molecules), device helpers and error checking.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace axtb {

constexpr int MAXPRIM = 6;
// status of a Jacobi job skipped by its run flag (neither converged nor failed)
constexpr int JACOBI_SKIPPED = -2147483647 - 1;
constexpr int MAXREF = 7;
constexpr int NFREQ = 23;

// One molecule of a batch. Offsets index the batch-wide arrays; matrices of a molecule are
// nao x nao column-major at `mat0` of every matrix pool, atom-pair buffers nat x nat at `pair0`,
// shell-pair buffers nsh x nsh at `shp0`.
struct Mol {
  int nat, nsh, nao, ndim;       // ndim = nsh + 9 nat: the SCC vector [q_sh | mu | theta]
  int at0, sh0, ao0, vec0;
  long long mat0, pair0, shp0, hist0;
  double nalpha, nbeta, kt, charge;
  double pair_cut2;              // atom pairs of the integrals within this squared distance (bohr^2)
};

// Per-atom parameters (SoA over the batch).
struct AtomP {
  const int* z;
  const int* sp;                 // compact species index of the D4 C6 table
  const int* mol;                // molecule of the atom
  const int* ish;                // first shell (batch index)
  const int* nsh;
  const double* xyz;             // bohr, [3*a]
  const double *rcov_cn, *en, *rad;             // GFN2 CN radius, Pauling EN, shell-polynomial radius
  const double *rep_alpha, *rep_zeff;           // repulsion
  const double *dkernel, *qkernel, *aes_rad, *aes_vcn;
  const double *eeq_chi, *eeq_eta, *eeq_kcnchi, *eeq_rad, *rcov_eeq;
  const double *d4_rcov, *d4_zeff, *d4_gi, *d4_r4r2, *d4_refcn, *d4_refq;  // refcn/refq [7*a]
  const int *d4_nref, *d4_ngw;                  // ngw [7*a]
};

// Per-shell parameters (SoA over the batch).
struct ShellP {
  const int *at, *l, *nprim, *iao;              // iao: first AO within the molecule
  const double *alpha, *coeff;                  // [MAXPRIM*s]
  const double *se0, *kcn, *shpoly, *refocc, *hard, *gam3, *minalpha, *zeta;
};

struct Batch {
  int nmol, nat, nsh, nao;
  const Mol* mol;
  AtomP a;
  ShellP s;
  const int *ao_sh, *ao_at;     // batch indices of each AO's shell and atom
  const double* c6ref;          // [nsp][nsp][7][7]
  int nsp;
  double intcut;
  // real-space cutoffs of the conventions: squared distances (bohr^2) of the GFN2 CN, the D4 CN and
  // the repulsion; radii and switch widths (bohr) of the D4 two-body and three-body terms
  double cn_cut2, d4cn_cut2, rep_cut2, d4_rc_2b, d4_w_2b, d4_rc_3b, d4_w_3b;
};

#define HIP_OK(x)                                                                                \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess)                                                                        \
      throw axtb::DeviceError(std::string(#x) + ": " + hipGetErrorString(e_));                    \
  } while (0)

struct DeviceError : std::runtime_error {
  explicit DeviceError(const std::string& m) : std::runtime_error(m) {}
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over a workgroup of `nthreads` (multiple of 64) threads; `red` holds nthreads/64 doubles.
// The result is returned to every thread; the summation order is fixed.
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  int nw = blockDim.x >> 6;
  for (int i = 0; i < nw; i++) s += red[i];
  return s;
}

__device__ __forceinline__ double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

}  // namespace axtb
