// GFN2-xTB energy + analytic gradient on one GPU: host side of the engine and its C ABI.
//
// A calculation takes a batch of molecules (one molecule is a batch of one) through
//   setup    coordination numbers, EEQ start charges, gamma, repulsion, D4 (three-body at q = 0),
//            the S/H0/D/Q integral build and the Loewdin orthogonaliser X = S^-1/2;
//   SCC      potentials -> Fock update -> A = X F X -> eigenpairs -> C = X U -> Fermi filling ->
//            P = C f C^T -> Mulliken moments -> energy -> convergence test -> Broyden mixing,
//            every kernel over the molecules still iterating (per-molecule masks);
//   gradient W = C f eps C^T and the pair/integral-derivative kernels with both CN chain rules.
// The stages follow XtbModel::singlepoint_state (crates/axtb/src/xtb.rs:437-861) in axtb's
// default conventions (crates/axtb/src/conventions.rs:244-268).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>

#include "../../include/axtb_hip.h"
#include "gfn2_tables.h"
#include "kernels.hip"
#include "linalg.hip"
#include "dense_eigh.hip"
#include "solvation.hip"

namespace axtb {

namespace {

const double AATOAU_2022 = 1.0 / (5.29177210544e-11 * 1e10);  // base/src/units.rs:18-20
const double KB_HARTREE = 3.1668115634564e-6;                   // base/src/units.rs:24
const double KB_XTB_TBLITE = 3.166808578545117e-06;             // base/src/units.rs:81 (tblite's kB)
const double PI = 3.14159265358979323846;

struct Fail {
  int code;
  std::string msg;
};

// the vector mixer of one SCF attempt (MixerKind, scf/mixing.rs:83-118)
enum { MIX_BROYDEN = 0, MIX_DIIS = 1 };

// STO-nG contraction with tblite's primitive normalisation (basis.rs:20-81, math.rs:108-111)
void sto_ng(int ng, int n, int l, double zeta, double* alpha, double* coeff) {
  int ityp = (l == 0 ? n : (l == 1 ? 4 + n : 7 + n)) - 1;
  const double DFACT[3] = {1.0, 1.0, 3.0};  // (2l - 1)!!
  double z2 = zeta * zeta;
  for (int i = 0; i < ng; i++) {
    double a, c;
    switch (ng) {
      case 1: a = STO_P_ALPHA1[ityp]; c = 1.0; break;
      case 2: a = STO_P_ALPHA2[ityp][i]; c = STO_P_COEFF2[ityp][i]; break;
      case 3: a = STO_P_ALPHA3[ityp][i]; c = STO_P_COEFF3[ityp][i]; break;
      case 4: a = STO_P_ALPHA4[ityp][i]; c = STO_P_COEFF4[ityp][i]; break;
      case 5: a = STO_P_ALPHA5[ityp][i]; c = STO_P_COEFF5[ityp][i]; break;
      default:  // six primitives; n = 6 has its own 6s / 6p expansions (basis.rs:56-61)
        a = n == 6 ? (l == 0 ? STO_P_ALPHA_6S[i] : STO_P_ALPHA_6P[i]) : STO_P_ALPHA6[ityp][i];
        c = n == 6 ? (l == 0 ? STO_P_COEFF_6S[i] : STO_P_COEFF_6P[i]) : STO_P_COEFF6[ityp][i];
        break;
    }
    alpha[i] = a * z2;
    double top = 2.0 / PI;
    coeff[i] = c * (std::pow(top * alpha[i], 0.75) * std::pow(std::sqrt(4.0 * alpha[i]), l) / std::sqrt(DFACT[l]));
  }
}

// Whether the GFN2 basis of element z has shells up to d only (what the integral kernels take:
// every element of the parametrisation, H-Rn).
bool supported_element(int z) {
  if (z < 1 || z > AXTB_GFN2_MAX_ELEM) return false;
  for (int s = 0; s < GFN2_NSHELL[z - 1]; s++)
    if (GFN2_ANG_SHELL[z - 1][s] > 2) return false;
  return true;
}

// Pinned host staging for the small per-launch uploads (job lists): a copy from pageable memory
// makes the host wait for the stream to drain, one from pinned memory does not. A ring; when it
// wraps, the device is synchronised so that no copy still reads the part being reused.
// the arena of the no-allocation call running on this host thread (null: buffers use hipMalloc)
struct Arena;
thread_local Arena* t_arena = nullptr;

struct Staging {
  char* base = nullptr;
  size_t cap = 0, off = 0;
  const void* put(const void* src, size_t bytes) {
    if (t_arena) return nullptr;  // no host allocation in a no-allocation call: pageable copies
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (!base) {
      cap = (size_t)64 << 20;
      if (hipHostMalloc((void**)&base, cap, hipHostMallocDefault) != hipSuccess) {
        base = nullptr;
        cap = 0;
      }
    }
    if (!base || need > cap) return nullptr;  // the caller copies from pageable memory
    if (off + need > cap) {
      HIP_OK(hipDeviceSynchronize());
      off = 0;
    }
    char* d = base + off;
    off += need;
    std::memcpy(d, src, bytes);
    return d;
  }
  ~Staging() {
    if (base) {
      (void)hipDeviceSynchronize();
      (void)hipHostFree(base);
    }
  }
};

// Device memory handed out from a fixed region instead of hipMalloc (the no-allocation entry
// point): a bump allocator over the region; reset() makes every buffer it served forget its memory,
// so that the next call carves them again from the start.
struct DBufBase {
  virtual void forget() = 0;
  virtual ~DBufBase() {}
};
struct Arena {
  char* base = nullptr;
  size_t cap = 0, off = 0, high = 0;
  std::vector<DBufBase*> users;
  void* carve(size_t bytes, DBufBase* u) {
    const size_t need = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
    if (off + need > cap)
      throw Fail{AXTB_ERR_INPUT, "the calculation needs more device memory than the no-allocation scratch holds (" +
                                     std::to_string(cap >> 20) + " MiB)"};
    void* p = base + off;
    off += need;
    high = std::max(high, off);
    users.push_back(u);
    return p;
  }
  void drop(DBufBase* u) { users.erase(std::remove(users.begin(), users.end(), u), users.end()); }
  void reset() {
    for (DBufBase* u : users) u->forget();
    users.clear();
    off = 0;
  }
};

template <class T>
struct DBuf : DBufBase {
  T* p = nullptr;
  size_t n = 0;
  Arena* arena = nullptr;  // the region p was carved from (null: hipMalloc)
  void forget() override {
    p = nullptr;
    n = 0;
    arena = nullptr;
  }
  void ensure(size_t m) {
    if (m <= n) return;
    if (t_arena) {
      if (p && !arena) throw DeviceError("a heap buffer in the no-allocation engine");
      p = (T*)t_arena->carve(std::max<size_t>(m, 1) * sizeof(T), this);
      arena = t_arena;
      n = std::max<size_t>(m, 1);
      return;
    }
    if (p && arena) throw DeviceError("an arena buffer outside a no-allocation call");
    if (p) HIP_OK(hipFree(p));
    p = nullptr;
    HIP_OK(hipMalloc(&p, std::max<size_t>(m, 1) * sizeof(T)));
    n = std::max<size_t>(m, 1);
  }
  // grow to `m` elements keeping the first `keep` (copied on stream s)
  void grow_keep(size_t keep, size_t m, hipStream_t s) {
    if (m <= n) return;
    T* old = p;
    const bool old_heap = p && !arena;
    p = nullptr;
    n = 0;
    arena = nullptr;
    if (t_arena) {
      if (old_heap) throw DeviceError("a heap buffer in the no-allocation engine");
      p = (T*)t_arena->carve(m * sizeof(T), this);
      arena = t_arena;
    } else {
      HIP_OK(hipMalloc(&p, m * sizeof(T)));
    }
    n = m;
    if (keep && old) HIP_OK(hipMemcpyAsync(p, old, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
    if (old_heap) {
      HIP_OK(hipStreamSynchronize(s));
      HIP_OK(hipFree(old));
    }
  }
  void upload(const std::vector<T>& v, hipStream_t s, Staging* stage = nullptr) {
    ensure(v.size());
    if (v.empty()) return;
    const void* src = stage ? stage->put(v.data(), v.size() * sizeof(T)) : nullptr;
    HIP_OK(hipMemcpyAsync(p, src ? src : (const void*)v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
  ~DBuf() {
    if (arena) {
      arena->drop(this);
    } else if (p) {
      (void)hipFree(p);
    }
  }
};

inline unsigned blocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// Kernel timing with HIP events on the engine's stream (enabled per context): every launch
// through LAUNCH records a start/stop pair under the kernel's name.
struct Prof {
  bool on = false;
  struct Rec {
    std::string name;
    hipEvent_t a, b;
    double work;  // algorithmic flops (GEMM, eigensolver) or HBM bytes (streaming kernels) of the launch
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  hipEvent_t ev() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIP_OK(hipEventCreate(&e));
    return e;
  }
  struct Sum {
    long n = 0;
    double ms = 0, work = 0;
    double span = 0;  // ms covered by the union of the launches' intervals (launches on several streams overlap)
  };
  // host-side counters reported like kernels (launches = items, work = total), 0 ms
  std::map<std::string, Sum> counters;
  void count(const std::string& name, long items, double total) {
    auto& x = counters[name];
    x.n += items;
    x.work += total;
  }
  // name -> launches, total ms, total work; the recorded events are returned to the pool
  std::map<std::string, Sum> drain() {
    std::map<std::string, Sum> m = counters;
    counters.clear();
    std::map<std::string, std::vector<std::pair<double, double>>> iv;
    for (auto& r : recs) {
      HIP_OK(hipEventSynchronize(r.b));
      float ms = 0.f, t0 = 0.f;
      HIP_OK(hipEventElapsedTime(&ms, r.a, r.b));
      HIP_OK(hipEventElapsedTime(&t0, recs[0].a, r.a));  // start relative to the first launch
      auto& x = m[r.name];
      x.n++;
      x.ms += ms;
      x.work += r.work;
      iv[r.name].push_back({(double)t0, (double)t0 + (double)ms});  // an interval's length is its ms exactly
    }
    for (auto& kv : iv) {
      auto& v = kv.second;
      std::sort(v.begin(), v.end());
      double span = 0.0, a = v[0].first, b = v[0].second;
      for (size_t k = 1; k < v.size(); k++) {
        if (v[k].first > b) {
          span += b - a;
          a = v[k].first;
          b = v[k].second;
        } else {
          b = std::max(b, v[k].second);
        }
      }
      m[kv.first].span = span + (b - a);
    }
    for (auto& r : recs) {
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    recs.clear();
    return m;
  }
};

struct ProfScope {
  Prof* p;
  hipStream_t s;
  Prof::Rec r;
  ProfScope(Prof* p_, hipStream_t s_, const char* name, double work = 0.0) : p(p_), s(s_) {
    if (p->on) {
      r.name = name;
      r.work = work;
      r.a = p->ev();
      r.b = p->ev();
      HIP_OK(hipEventRecord(r.a, s));
    }
  }
  ~ProfScope() {
    if (p->on) {
      (void)hipEventRecord(r.b, s);
      p->recs.push_back(r);
    }
  }
};

#define LAUNCH(kern, grid, block, shm, ...) LAUNCHW(0.0, kern, grid, block, shm, __VA_ARGS__)
#define LAUNCHS(stream, work, kern, grid, block, shm, ...)                  \
  do {                                                                      \
    ProfScope ps_(&prof, stream, #kern, work);                              \
    hipLaunchKernelGGL(kern, grid, block, shm, stream, __VA_ARGS__);        \
  } while (0)
#define LAUNCHW(work, kern, grid, block, shm, ...)                          \
  do {                                                                      \
    ProfScope ps_(&prof, st, #kern, work);                                  \
    hipLaunchKernelGGL(kern, grid, block, shm, st, __VA_ARGS__);            \
  } while (0)

struct MolIn {
  int nat;
  const int* z;
  const double* xyz;
  double charge;
  int uhf;
  // a given SCF start (xtb.rs:204-208): shell charges [nsh] (null: none), atomic dipoles
  // [nat][3] and quadrupoles [nat][6] (null: zero)
  const double* gq = nullptr;
  const double* gdp = nullptr;
  const double* gqp = nullptr;
};

struct MolOut {
  int status = 0;
  std::string msg;
  double energy = 0, efermi = 0, residual = 0, energy_change = 0;
  double terms[9] = {0};
  int iterations = 0, nsh = 0, nao = 0;
  std::vector<double> grad, qat, qsh, dp, qp, emo;
  double dipole[3] = {0, 0, 0}, gap = 0;
  // ScfReport: mixer of the result (0 broyden, 1 diis, 2 diis-fallback), attempts
  int mixer_used = 0, nattempts = 0, att_iter[2] = {0, 0}, att_conv[2] = {0, 0};
  int element = 0;  // the unsupported element
};

// the SCF attempts of one molecule (MixerPlan, scf/mixing.rs:120-170)
struct Attempts {
  int n = 0, kind[2] = {0, 0}, iter[2] = {0, 0}, conv[2] = {0, 0};
};

struct Engine {
  Staging stage;  // first member: destroyed last, after every buffer and stream
  int device = 0;
  hipStream_t st = nullptr;
  hipStream_t st2 = nullptr;  // side stream (default priority) for overlapped eigensolver launches
  hipStream_t st3 = nullptr;  // a second one (default priority) for the batched Jacobi's size classes
  hipStream_t st4 = nullptr;  // a third one (default priority): a batched Jacobi beside other work
  bool jacobi_streams = true;  // AXTB_HIP_JACOBI_STREAMS
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join3 = nullptr, ev_fork4 = nullptr, ev_join4 = nullptr;
  hipEvent_t ev_act[2] = {nullptr, nullptr};  // the active mask of an SCC iteration of either parity copied
  Prof prof;
  std::string last_error;
  // host mirrors of the batch tables
  std::vector<Mol> mols;
  // device tables
  DBuf<Mol> d_mol;
  DBuf<int> i_z, i_sp, i_mol, i_ish, i_nsh, i_d4nref, i_d4ngw, s_at, s_l, s_np, s_iao, ao_sh, ao_at;
  DBuf<double> a_xyz, a_rcov, a_en, a_rad, a_ralpha, a_rzeff, a_dk, a_qk, a_arad, a_avcn, a_chi, a_eta, a_kcn,
      a_erad, a_reeq, a_d4rcov, a_d4zeff, a_d4gi, a_d4r4r2, a_d4refcn, a_d4refq;
  DBuf<double> s_alpha, s_coeff, s_se0, s_kcn, s_shpoly, s_refocc, s_hard, s_gam3, s_minal, s_zeta, c6ref;
  // work
  DBuf<double> cn, cneeq, cnd4, se, dsedcn, mrad, dmrdcn, gam, pool, X, Bm, F, T, Am, Vw, Z, C, P, Wm, w, focc, fwork,
      fst, eeqA, eeqV, eeqZ, eeqw, eeqr, eeqy, qstart, gw, dgw, atmP, e3, g3, dcn3, erep, grep;
  DBuf<double> qin, qout, dq, qlast, dqlast, hdf, hu, hs, qsh, qat, dp, qp, qsh_o, qat_o, dp_o, qp_o, vsh, vat, vdp,
      vqp, gq, cols, eband, eel, eterm, mst, grad, dedcn, dedcn4;
  // Broyden history (slot-major, grown with the iterations: ensure_hist), DIIS ring and state
  DBuf<double> hom, ham, gbeta, dhq, dhr, dbas, dbq, dbr, ddst;
  DBuf<int> dist, dmask;
  int hist_slots = 0;
  // df and u of `slots` history slots (slot * tot_vec + vec0), their omega (slot * nmol + m) and
  // overlaps (packed triangle * nmol + m); a grown buffer keeps the slots already written
  void ensure_hist(int slots) {
    if (slots <= hist_slots) return;
    const int ns = std::min(cap, std::max(slots, std::max(16, 2 * hist_slots)));
    const size_t NM = mols.size();
    auto grow = [&](DBuf<double>& d, size_t keep, size_t need) { d.grow_keep(keep, need, st); };
    auto tri = [](size_t k) { return k * (k + 1) / 2; };
    grow(hdf, (size_t)hist_slots * tot_vec, (size_t)ns * tot_vec);
    grow(hu, (size_t)hist_slots * tot_vec, (size_t)ns * tot_vec);
    grow(hom, (size_t)hist_slots * NM, (size_t)ns * NM);
    grow(ham, tri(hist_slots) * NM, tri(ns) * NM);
    hist_slots = ns;
  }
  DBuf<int> active, status, kocc, itc, eeqfail;
  DBuf<long long> eq0;
  DBuf<GemmJob> gj;
  DBuf<EigJob> ej;

  // matrices of order above dense_min are diagonalised by the dense solver (Householder, divide
  // and conquer), smaller ones by the batched Jacobi solver; AXTB_HIP_DENSE_MIN overrides
  int dense_min = 200;
  // the tridiagonalisation's symv takes two tiles per workgroup from this many tiles of the
  // trailing matrix on (AXTB_HIP_TRD_PAIR_MIN)
  int trd_pair_min = 2048;

  explicit Engine(int dev) : device(dev) {
    HIP_OK(hipSetDevice(dev));
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&st3, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&st4, hipStreamNonBlocking));
    HIP_OK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&ev_join3, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&ev_fork4, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&ev_join4, hipEventDisableTiming));
    for (auto& e : ev_act) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // the batched Jacobi kernel handles n <= 200
    if (const char* v = std::getenv("AXTB_HIP_DENSE_MIN")) dense_min = std::min(200, std::max(1, std::atoi(v)));
    if (const char* v = std::getenv("AXTB_HIP_TRD_PAIR_MIN")) trd_pair_min = std::max(1, std::atoi(v));
    if (const char* v = std::getenv("AXTB_HIP_OA_STEPS")) oa_steps = std::min(4, std::max(1, std::atoi(v)));
    if (const char* v = std::getenv("AXTB_HIP_DENSE_FULL")) dense_full = std::max(1, std::atoi(v));
    if (const char* v = std::getenv("AXTB_HIP_SPLIT_REFINE")) split_refine = std::atoi(v) != 0;
    if (const char* v = std::getenv("AXTB_HIP_JACOBI_WIDE")) jacobi_wide = std::atoi(v) != 0;
    if (const char* v = std::getenv("AXTB_HIP_JACOBI_STREAMS")) jacobi_streams = std::atoi(v) != 0;
    if (const char* v = std::getenv("AXTB_HIP_SPLIT_RESIDUAL")) split_residual = std::atof(v);
    if (const char* v = std::getenv("AXTB_HIP_SPLIT_STEPS")) split_steps = std::max(3, std::atoi(v));
    if (const char* v = std::getenv("AXTB_HIP_SPLIT_FROM")) split_from = std::max(2, std::atoi(v));
    if (const char* v = std::getenv("AXTB_HIP_DENSE_REFINE")) dense_refine = std::atoi(v) != 0;
    if (const char* v = std::getenv("AXTB_HIP_GEMM_SMALL")) gemm_small = std::atoi(v) != 0;
    if (const char* v = std::getenv("AXTB_HIP_UPD32")) upd32 = std::atoi(v) != 0;
    if (const char* v = std::getenv("AXTB_HIP_ROCBLAS")) use_rocblas = std::atoi(v) != 0;
    if (const char* v = std::getenv("AXTB_HIP_ROCBLAS_MIN")) rocblas_min = std::max(1, std::atoi(v));
    if (const char* v = std::getenv("AXTB_HIP_EXACT_LEVELS")) exact_levels = std::atoi(v) != 0;
    if (const char* v = std::getenv("AXTB_HIP_DENSE_ACCEPT")) dense_accept = std::min(1e-3, std::max(1e-12, std::atof(v)));
    // a lower limit makes Jacobi solves end unconverged (the error path's test)
    if (const char* v = std::getenv("AXTB_HIP_JACOBI_SWEEPS")) jacobi_sweeps = std::min(60, std::max(1, std::atoi(v)));
  }
  bool dense(int n) const { return n > dense_min; }

  // All eigenpairs of one symmetric matrix by the dense solver: A (destroyed), w, Z (n x n);
  // Q is an n x n workspace.
  DBuf<double> dtau1;
  void dense_eigh(double* A, int n, double* w, double* Q, double* Z) {
    dtau1.ensure(n);
    dense_values(A, n, w, dtau1.p, Q);
    dense_vectors(A, n, dtau1.p, Q, n, Z);
  }
  ~Engine() {
    if (blas) (void)rocblas_destroy_handle(blas);
    if (st2) (void)hipStreamDestroy(st2);
    if (st3) (void)hipStreamDestroy(st3);
    if (st4) (void)hipStreamDestroy(st4);
    if (st) (void)hipStreamDestroy(st);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (ev_join3) (void)hipEventDestroy(ev_join3);
    if (ev_fork4) (void)hipEventDestroy(ev_fork4);
    if (ev_join4) (void)hipEventDestroy(ev_join4);
    for (auto& e : ev_act)
      if (e) (void)hipEventDestroy(e);
    if (pin_act) (void)hipHostFree(pin_act);
  }

  // dj: the same jobs already on the device (uploaded once by the caller for repeated launches);
  // they must all be non-empty and of one kind (FP32 updates or not), and are never sent to the
  // library or split over K (those paths cannot honour run flags)
  // what a launch needs to know of a job list, in one pass over it (kept for the fixed device lists,
  // which are launched in every SCC iteration)
  struct GemmMeta {
    int mm = 0, nn = 0;
    bool same = true, two = false, any32 = false, all32 = true, anyrun = false, uniform = true;
    double flops = 0.0;  // a lower-triangle (symmetric) output counts its m (m + 1) / 2 elements
  };
  static GemmMeta gemm_meta(const std::vector<GemmJob>& jobs) {
    GemmMeta M;
    for (const GemmJob& j : jobs) {
      M.mm = std::max(M.mm, j.m);
      M.nn = std::max(M.nn, j.n);
      M.same &= j.ta == jobs[0].ta && j.tb == jobs[0].tb;
      M.two |= j.k2 > 0;
      M.any32 |= j.f32upd != 0;
      M.all32 &= j.f32upd != 0;
      M.anyrun |= j.run != nullptr;
      M.uniform &= j.m > 0 && j.n > 0 && j.f32upd == jobs[0].f32upd;
      M.flops += (j.lower && j.m == j.n ? (double)j.m * (j.m + 1) : 2.0 * j.m * j.n) * ((double)j.k + j.k2);
    }
    return M;
  }
  std::vector<std::pair<std::pair<const GemmJob*, size_t>, GemmMeta>> gmeta;  // of the fixed device lists

  void gemm(std::vector<GemmJob>& jobs, const GemmJob* dj = nullptr) {
    GemmMeta M;
    if (dj) {
      bool found = false;
      for (const auto& e : gmeta)
        if (e.first.first == dj && e.first.second == jobs.size()) {
          M = e.second;
          found = true;
          break;
        }
      if (!found) {
        M = gemm_meta(jobs);
        if (!M.uniform) throw DeviceError("gemm: device job list not uniform");
        gmeta.push_back({{dj, jobs.size()}, M});
      }
    } else {
      jobs.erase(std::remove_if(jobs.begin(), jobs.end(), [](const GemmJob& j) { return j.m <= 0 || j.n <= 0; }),
                 jobs.end());
      M = gemm_meta(jobs);
    }
    if (jobs.empty()) return;
    const bool flagged = dj || M.anyrun;
    auto jobs_dev = [&]() -> const GemmJob* {
      if (dj) return dj;
      gj.upload(jobs, st, &stage);
      return gj.p;
    };
    const bool any32 = M.any32;
    if (any32 && !M.all32) {
      std::vector<GemmJob> j32, j64;
      for (auto& j : jobs) (j.f32upd ? j32 : j64).push_back(j);
      gemm(j32);
      gemm(j64);
      return;
    }
    // one large plain product (no inner scaling, no second term, the full C): rocBLAS's dgemm, which
    // runs at a higher fraction of the FP64 matrix peak than the 128-tile kernel at these sizes (not in
    // a no-allocation call: the library keeps its own workspace). Symmetric products stay on the
    // lower tiles of the 128-tile kernel: the library's full product took longer (6.6 against 6 ms at
    // n = 6000).
    if (!any32 && use_rocblas && !t_arena && jobs.size() == 1 && !flagged) {
      const GemmJob& J = jobs[0];
      if (!J.dk && !J.A2 && !J.lower && J.m >= rocblas_min && J.n >= rocblas_min && J.k >= rocblas_min) {
        if (!blas) {
          if (rocblas_create_handle(&blas) != rocblas_status_success) throw DeviceError("rocblas_create_handle failed");
        }
        if (rocblas_set_stream(blas, st) != rocblas_status_success) throw DeviceError("rocblas_set_stream failed");
        const double al = J.alpha, be = J.beta;
        ProfScope ps_(&prof, st, "rocblas_dgemm", 2.0 * J.m * J.n * (double)J.k);
        const rocblas_status rs =
            rocblas_dgemm(blas, J.ta ? rocblas_operation_transpose : rocblas_operation_none,
                          J.tb ? rocblas_operation_transpose : rocblas_operation_none, J.m, J.n, J.k, &al, J.A,
                          (rocblas_int)J.lda, J.B, (rocblas_int)J.ldb, &be, J.C, (rocblas_int)J.ldc);
        if (rs != rocblas_status_success) throw DeviceError("rocblas_dgemm failed: " + std::to_string((int)rs));
        return;
      }
    }
    if (any32 && use_rocblas && !t_arena && jobs.size() == 1 && !flagged && jobs[0].m >= rocblas_min &&
        jobs[0].n >= rocblas_min && !jobs[0].ta && !jobs[0].tb && jobs[0].k == jobs[0].n && !jobs[0].A2 &&
        !jobs[0].dk) {
      // one large update: U and E = B - I rounded to FP32, U E by rocBLAS's sgemm (FP32 accumulation,
      // a higher fraction of the FP32 matrix peak than the 128-tile kernel), C = U + U E in FP64
      const GemmJob& J = jobs[0];
      const long long mk = (long long)J.m * J.k, kn = (long long)J.k * J.n, mn = (long long)J.m * J.n;
      f32buf.ensure((size_t)(mk + kn + mn));
      float *Uf = f32buf.p, *Ef = f32buf.p + mk, *Pf = f32buf.p + mk + kn;
      LAUNCH(k_to_f32, dim3(blocks(mk, 256)), dim3(256), 0, J.A, (long long)J.lda, J.m, J.k, 0, Uf);
      LAUNCH(k_to_f32, dim3(blocks(kn, 256)), dim3(256), 0, J.B, (long long)J.ldb, J.k, J.n, 1, Ef);
      if (!blas) {
        if (rocblas_create_handle(&blas) != rocblas_status_success) throw DeviceError("rocblas_create_handle failed");
      }
      if (rocblas_set_stream(blas, st) != rocblas_status_success) throw DeviceError("rocblas_set_stream failed");
      const float one = 1.0f, zero = 0.0f;
      {
        ProfScope ps_(&prof, st, "rocblas_sgemm_upd32", 2.0 * J.m * J.n * (double)J.k);
        const rocblas_status rs = rocblas_sgemm(blas, rocblas_operation_none, rocblas_operation_none, J.m, J.n, J.k,
                                                &one, Uf, J.m, Ef, J.k, &zero, Pf, J.m);
        if (rs != rocblas_status_success) throw DeviceError("rocblas_sgemm failed: " + std::to_string((int)rs));
      }
      LAUNCH(k_add_f32, dim3(blocks(mn, 256)), dim3(256), 0, J.A, (long long)J.lda, (const float*)Pf, J.m, J.n, J.C,
             (long long)J.ldc);
      HIP_OK(hipGetLastError());
      return;
    }
    if (any32) {  // refinement updates C = A + A (B - I), the product in FP32
      int mm = 0, nn = 0;
      double flops = 0.0;
      for (auto& j : jobs) {
        if (!j.f32upd || j.ta || j.tb || j.k != j.n || j.m < 1 || j.A2 || j.dk)
          throw DeviceError("gemm: an FP32 update is C = A + A (B - I) with square B");
        mm = std::max(mm, j.m);
        nn = std::max(nn, j.n);
        flops += 2.0 * j.m * j.n * (double)j.k;
      }
      const GemmJob* jp = jobs_dev();
      LAUNCHW(flops, k_gemm128_upd32, dim3(blocks(nn, BT), blocks(mm, BT), (unsigned)jobs.size()), dim3(256), 0, jp);
      HIP_OK(hipGetLastError());
      return;
    }
    const int mm = M.mm, nn = M.nn;
    const bool same = M.same, two = M.two;
    const double flops = M.flops;
    // one product with few output tiles and a long inner dimension: split K over workgroups,
    // partials reduced in a fixed order
    if (jobs.size() == 1 && !two && !flagged) {
      const GemmJob& J = jobs[0];
      const long long tiles = (long long)blocks(J.m, BT) * blocks(J.n, BT);
      const int splits = (int)std::min<long long>(std::min<long long>(256 / std::max(1LL, tiles), J.k / 256), 64);
      if (tiles < 128 && splits >= 2) {
        const int kc = (J.k + splits - 1) / splits;
        std::vector<GemmJob> parts;
        gpart.ensure((size_t)splits * J.m * J.n);
        for (int q = 0; q < splits; q++) {
          const int k0 = q * kc, k1 = std::min(J.k, k0 + kc);
          if (k1 <= k0) break;
          GemmJob P = J;
          P.k = k1 - k0;
          P.A = J.ta == 0 ? J.A + (long long)k0 * J.lda : J.A + k0;
          P.B = J.tb == 0 ? J.B + k0 : J.B + (long long)k0 * J.ldb;
          P.dk = J.dk ? J.dk + k0 : nullptr;
          P.C = gpart.p + (size_t)parts.size() * J.m * J.n;
          P.ldc = J.m;
          P.beta = 0.0;
          parts.push_back(P);
        }
        gj.upload(parts, st, &stage);
        dim3 grid(blocks(J.n, BT), blocks(J.m, BT), (unsigned)parts.size());
        launch128(J.ta * 2 + J.tb, grid, flops, gj.p);
        LAUNCH(k_gemm_reduce, dim3(blocks((long long)J.m * J.n, 256)), dim3(256), 0, (const double*)gpart.p,
               (int)parts.size(), J.m, J.n, J.C, J.ldc, J.beta);
        HIP_OK(hipGetLastError());
        return;
      }
    }
    // products with m, n <= 128: one workgroup per job, 16 x 16 granularity
    if (gemm_small && same && !two && mm <= 128 && nn <= 128 && tune_gemm_small[tune_bucket(std::max(mm, nn))] != 2) {
      const GemmJob* jp = jobs_dev();
      const int t = jobs[0].ta * 2 + jobs[0].tb;
      const dim3 grid((unsigned)jobs.size());
      if (t == 0)
        LAUNCHW(flops, k_gemm_small_nn, grid, dim3(512), 0, jp);
      else if (t == 1)
        LAUNCHW(flops, k_gemm_small_nt, grid, dim3(512), 0, jp);
      else if (t == 2)
        LAUNCHW(flops, k_gemm_small_tn, grid, dim3(512), 0, jp);
      else
        LAUNCHW(flops, k_gemm_small_tt, grid, dim3(512), 0, jp);
      HIP_OK(hipGetLastError());
      return;
    }
    // large products on 128 x 128 tiles (one transpose pair per launch; two-term products always),
    // the other batches of per-molecule products on 64 x 64 tiles
    if (two || (same && (long long)mm * nn >= 256LL * 256 && jobs.size() <= 8) ||
        (same && in_refine && std::max(mm, nn) > 64)) {
      if (!same) throw DeviceError("gemm: mixed transposes in one launch");
      const GemmJob* jp = jobs_dev();
      dim3 grid(blocks(nn, BT), blocks(mm, BT), (unsigned)jobs.size());
      if (two) {
        if (jobs[0].ta != 0 || jobs[0].tb != 1) throw DeviceError("gemm: two-term products are A B^T + A2 B2^T");
        for (auto& j : jobs)
          if (j.dk) throw DeviceError("gemm: two-term products take no inner scaling");
        LAUNCHW(flops, k_gemm128_nt2, grid, dim3(256), 0, jp);
      } else {
        launch128(jobs[0].ta * 2 + jobs[0].tb, grid, flops, jp);
      }
      HIP_OK(hipGetLastError());
      return;
    }
    const GemmJob* jp = jobs_dev();
    dim3 grid(blocks(nn, GT), blocks(mm, GT), (unsigned)jobs.size());
    LAUNCHW(flops, k_gemm, grid, dim3(256), 0, jp);
    HIP_OK(hipGetLastError());
  }
  DBuf<double> gpart;
  DBuf<float> f32buf;  // FP32 operands and product of a library FP32 update
  void launch128(int t, dim3 grid, double flops, const GemmJob* jp) {
    if (t == 0)
      LAUNCHW(flops, k_gemm128_nn, grid, dim3(256), 0, jp);
    else if (t == 1)
      LAUNCHW(flops, k_gemm128_nt, grid, dim3(256), 0, jp);
    else if (t == 2)
      LAUNCHW(flops, k_gemm128_tn, grid, dim3(256), 0, jp);
    else
      LAUNCHW(flops, k_gemm128_tt, grid, dim3(256), 0, jp);
  }

  // eigenpairs of every job's symmetric matrix (Jacobi, one workgroup per matrix)
  // work: the algorithmic flops of a dense eigendecomposition, (4/3) n^3 for the reduction to
  // tridiagonal form plus 2 n^2 k for the k needed eigenvectors (k = n here). Matrices that fit
  // the LDS of a workgroup run from LDS, the others from global memory (separate launches).
  int max_lds = 0;
  // A batched Jacobi solve as a plan: the jobs in launch order and the launches, uploaded once.
  // Launches: matrices that fit in LDS with their vectors (mode 2), with A only (mode 1), neither
  // (mode 0). Within a launch the largest matrices go first (shorter tail); the mode 1/0 launches
  // run on a second stream beside the mode 2 launch, both joined back into the engine stream.
  // fp32: the rotations in FP32 (A and V in LDS; Z and w returned in FP64 for the refinement).
  // tol_scale multiplies the rotation thresholds (the mixed path's FP32 solves, refined afterwards)
  // mol_of_job: the molecule of each job; a job that does not converge within max_sweeps then
  // marks its molecule with status bit 16 (an eigensolver failure, Error::Numerical). A plan over
  // jobs with run flags is launched as it is in every SCC iteration (the flags pick the jobs).
  struct EighPlan {
    struct Launch {
      int mode;
      size_t b0, b1, shm;
      unsigned nthr;
    };
    std::vector<Launch> launches;
    size_t moff[3] = {0, 0, 0}, mcnt[3] = {0, 0, 0};
    size_t njobs = 0;
    bool fp32 = false, flag = false, side = false;
    int max_sweeps = 60;
    float tol_scale = 1.0f;
    DBuf<EigJob> jobs;  // launch order: mode 2, then 1, then 0
    DBuf<int> stat, mol;
  };
  EighPlan eadhoc;  // the plan of eigh()

  void eigh_plan(EighPlan& P, const std::vector<EigJob>& jobs, bool fp32, int max_sweeps, float tol_scale,
                 const std::vector<int>* mol_of_job) {
    if (max_sweeps < 0) max_sweeps = jacobi_sweeps;
    if (!max_lds) {
      HIP_OK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
      if (max_lds <= 0) max_lds = 64 * 1024;
    }
    const size_t es = fp32 ? 4 : 8;
    std::vector<int> gidx[3];
    for (size_t q = 0; q < jobs.size(); q++) {
      const EigJob& j = jobs[q];
      if (j.n > 200) throw DeviceError("Jacobi eigensolver called for n > 200");
      // eigenvalues only (Z null): no V, so A alone in LDS (mode 1) when it fits; the tuning table
      // may keep V in HBM (mode 1) where both fit
      int mode = (j.Z && jacobi_lds(j.n, 2, es) <= (size_t)max_lds) ? 2 : (jacobi_lds(j.n, 1, es) <= (size_t)max_lds ? 1 : 0);
      if (mode == 2 && tune_jac_mode[fp32 ? 1 : 0][tune_bucket(j.n)] == 1) mode = 1;
      if (fp32 && mode == 0) throw DeviceError("FP32 Jacobi: matrix does not fit in LDS");
      gidx[mode].push_back((int)q);
    }
    std::vector<EigJob> order;
    std::vector<int> jm;  // the molecule of each job in launch order (the order of stat)
    for (int g = 0; g < 3; g++)
      std::stable_sort(gidx[g].begin(), gidx[g].end(), [&](int a, int b) { return jobs[a].n > jobs[b].n; });
    P.launches.clear();
    for (int mode = 2; mode >= 0; mode--) {
      P.moff[mode] = order.size();
      P.mcnt[mode] = gidx[mode].size();
      for (int q : gidx[mode]) {
        order.push_back(jobs[q]);
        if (mol_of_job) jm.push_back((*mol_of_job)[q]);
      }
      // the jobs (largest first) in launches of one LDS size class each: a launch reserves the LDS
      // of its largest matrix in every workgroup, so small matrices in a launch of their own fit 2-4
      // workgroups per CU
      auto per_cu = [&](int n) {
        return std::max<size_t>(1, std::min<size_t>(4, (size_t)max_lds / std::max<size_t>(jacobi_lds(n, mode, es), 1)));
      };
      const size_t e0 = P.moff[mode], e1 = order.size();
      size_t b0 = e0;
      while (b0 < e1) {
        const size_t shm = jacobi_lds(order[b0].n, mode, es);
        const size_t pc = mode == 0 ? 1 : per_cu(order[b0].n);
        size_t b1 = b0 + 1;
        while (b1 < e1 && (mode == 0 || per_cu(order[b1].n) <= pc)) b1++;
        // 16 wavefronts per matrix while they fit (32 per CU: one workgroup per CU, or two of the FP32
        // kernels, which use at most 64 registers), else 8
        unsigned nthr = (mode != 0 && jacobi_wide && (pc == 1 || (pc == 2 && fp32))) ? 1024u : 512u;
        const int tt = tune_jac_threads[fp32 ? 1 : 0][mode][tune_bucket(order[b0].n)];
        if (tt && jacobi_threads_ok(order[b0].n, tt)) nthr = (unsigned)tt;
        P.launches.push_back(EighPlan::Launch{mode, b0, b1, shm, nthr});
        b0 = b1;
      }
    }
    P.njobs = order.size();
    P.fp32 = fp32;
    P.flag = mol_of_job != nullptr;
    P.max_sweeps = max_sweeps;
    P.tol_scale = tol_scale;
    P.side = P.mcnt[2] && (P.mcnt[1] || P.mcnt[0]);
    P.jobs.upload(order, st, &stage);
    P.stat.ensure(order.size());
    if (P.flag) P.mol.upload(jm, st, &stage);
  }

  // Streams: the mode 1/0 launches on st2 beside the mode 2 ones; with jacobi_streams the mode 2
  // size classes after the first (largest) run on st3, so that they fill the CUs the first launch
  // leaves idle in its last round of workgroups.
  // base: the stream the solve is ordered on (the engine stream when null).
  void eigh_launch(EighPlan& P, hipStream_t base = nullptr) {
    if (!P.njobs) return;
    if (!base) base = st;
    int n2 = 0;
    for (const auto& L : P.launches) n2 += L.mode == 2;
    const bool split2 = jacobi_streams && n2 >= 2;
    if (P.side || split2) {
      HIP_OK(hipEventRecord(ev_fork, base));
      if (P.side) HIP_OK(hipStreamWaitEvent(st2, ev_fork, 0));
      if (split2) HIP_OK(hipStreamWaitEvent(st3, ev_fork, 0));
    }
    for (int mode = 2; mode >= 0; mode--) {
      if (!P.mcnt[mode]) continue;
      hipStream_t s = (P.side && mode < 2) ? st2 : base;
      const char* nm = P.fp32 ? (mode == 2 ? "k_jacobi2_f32" : "k_jacobi1_f32")
                              : (mode == 2 ? "k_jacobi2" : (mode == 1 ? "k_jacobi1" : "k_jacobi0"));
      const JacobiKernel kern = jacobi_kernel(mode, P.fp32);
      int nl = 0;
      for (const auto& L : P.launches) {
        if (L.mode != mode) continue;
        hipStream_t sl = (mode == 2 && split2 && nl++ > 0) ? st3 : s;
        // the work of the solves that ran is summed on the device (k_sweep_acc) when profiling
        ProfScope ps_(&prof, sl, nm, 0.0);
        HIP_OK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.shm));
        hipLaunchKernelGGL(kern, dim3((unsigned)(L.b1 - L.b0)), dim3(L.nthr), L.shm, sl, P.jobs.p + L.b0, P.max_sweeps,
                           P.stat.p + L.b0, P.tol_scale);
      }
      if (mode == 2 && split2) {
        HIP_OK(hipEventRecord(ev_join3, st3));
        HIP_OK(hipStreamWaitEvent(base, ev_join3, 0));
      }
      HIP_OK(hipGetLastError());
      if (prof.on) {
        // sweep statistics and the algorithmic work of the solves that ran, summed on the device
        // (read at the profile report)
        if (!sweep_acc.p) {
          sweep_acc.ensure(24);
          HIP_OK(hipMemset(sweep_acc.p, 0, 24 * sizeof(double)));
        }
        const int t = mode * 2 + (P.fp32 ? 1 : 0);
        hipLaunchKernelGGL(k_sweep_acc, dim3(1), dim3(256), 0, s, (const int*)P.stat.p + P.moff[mode],
                           (const EigJob*)P.jobs.p + P.moff[mode], (int)P.mcnt[mode], sweep_acc.p + 3 * t,
                           sweep_acc.p + 18 + t);
      }
    }
    if (P.side) {
      HIP_OK(hipEventRecord(ev_join, st2));
      HIP_OK(hipStreamWaitEvent(base, ev_join, 0));
    }
    if (P.flag)
      LAUNCHS(base, 0.0, k_jacobi_flag, dim3(blocks((long long)P.njobs, 256)), dim3(256), 0, (const int*)P.stat.p,
              (const int*)P.mol.p, (int)P.njobs, status.p);
  }

  void eigh(std::vector<EigJob>& jobs, bool fp32 = false, int max_sweeps = -1, float tol_scale = 1.0f,
            const std::vector<int>* mol_of_job = nullptr) {
    if (jobs.empty()) return;
    eigh_plan(eadhoc, jobs, fp32, max_sweeps, tol_scale, mol_of_job);
    eigh_launch(eadhoc);
  }
  // launch shapes per size bucket (n / 8 for n <= 207), set by the autotuner through
  // axtbhip_tuning_set; 0 keeps the built-in rule
  static constexpr int TUNE_NB = 26;
  int tune_jac_threads[2][3][TUNE_NB] = {};  // [FP32][mode][bucket]: 256, 512 or 1024 threads
  int tune_jac_mode[2][TUNE_NB] = {};         // [FP32][bucket]: 1 keeps V in HBM where mode 2 fits
  int tune_gemm_small[TUNE_NB] = {};          // [bucket of max(m, n)]: 1 one-workgroup kernel, 2 64-tile kernel
  static int tune_bucket(int n) { return std::min(TUNE_NB - 1, std::max(0, n / 8)); }
  // a Jacobi workgroup of t threads takes every 2 x 2 block of an n x n matrix (at most 10 per
  // thread) and one rotation pair per thread
  static bool jacobi_threads_ok(int n, int t) {
    const int np = (n + (n & 1)) / 2;
    return (t == 256 || t == 512 || t == 1024) && (long long)t * 10 >= (long long)np * (np + 1) / 2 && t > np;
  }
  DBuf<double> sweep_acc;  // per Jacobi mode and precision: sweeps, solves, unconverged solves; then the work
  // the device-side sweep statistics into the profile's counters
  void drain_sweeps() {
    if (rd_cnt.p) {  // the device-driven refinements: [full, split] x [jobs, steps, failures]
      unsigned long long c[6];
      HIP_OK(hipMemcpy(c, rd_cnt.p, sizeof c, hipMemcpyDeviceToHost));
      for (int k = 0; k < 2; k++) {
        if (c[3 * k]) prof.count(k ? "refine_steps_split" : "refine_steps", (long)c[3 * k], (double)c[3 * k + 1]);
        if (c[3 * k + 2])
          prof.count(k ? "refine_split_fallbacks" : "refine_fallbacks", (long)c[3 * k + 2], (double)c[3 * k + 2]);
      }
      HIP_OK(hipMemset(rd_cnt.p, 0, sizeof c));
    }
    if (!sweep_acc.p) return;
    double h[24];
    HIP_OK(hipMemcpy(h, sweep_acc.p, sizeof h, hipMemcpyDeviceToHost));
    for (int t = 0; t < 6; t++) {
      const std::string tag = std::to_string(t / 2) + (t % 2 ? "_f32" : "");
      if (h[3 * t + 1] > 0) prof.count("sweeps_jacobi" + tag, (long)h[3 * t + 1], h[3 * t]);
      if (h[3 * t + 2] > 0) prof.count("unconverged_jacobi" + tag, (long)h[3 * t + 2], h[3 * t + 2]);
      if (h[18 + t] > 0) prof.count("k_jacobi" + tag, 0, h[18 + t]);  // work only: launches come from the events
    }
    HIP_OK(hipMemset(sweep_acc.p, 0, sizeof h));
  }
  int jacobi_sweeps = 60;  // sweeps of a Jacobi solve before it counts as not converged

  // ---------------------------------------------------------------- dense eigensolver (large n)
  // Eigenvalues of the symmetric n x n matrix A into w (ascending). A is overwritten with the
  // Householder vectors, tau (n) receives their scalars and Q (n x n) the eigenvectors of the
  // tridiagonal matrix: the state dense_vectors() transforms back.
  DBuf<double> dtrW, dtrDots, dtrVp, dtrSq, dtrAx;
  DBuf<double> dd, de, dpart, dq_Q2, dD, dD2, dds, dzs, drot_cs, dkd, dkz, dtaus, dzh, dval, dU, dG, dQG, dleaf,
      dleafV, dleafZ, dleafw, dbtV, dbtY, dbtVtV, dbtVT;
  DBuf<int> dperm, dkind, drot_ij, dkcol, dorg, dsrc, dnk, dnr, dls, dlm;
  DBuf<double> drho;
  DBuf<long long> duoff, dgoff;
  DBuf<DcNode> dnodes;

  void dense_values(double* A, int n, double* w, double* tau, double* Q) {
    for (auto* b : {&dd, &de, &dD, &dD2, &dds, &dzs, &dkd, &dkz, &dtaus, &dzh, &dval}) b->ensure(n);
    drot_cs.ensure(2 * (size_t)n);
    for (auto* b : {&dperm, &dkind, &dkcol, &dorg, &dsrc}) b->ensure(n);
    drot_ij.ensure(2 * (size_t)n);
    // 1. blocked Householder tridiagonalisation
    const int nchmax = (n + TRD_CH - 1) / TRD_CH;
    const int ndbmax = (n + TRD_DR - 1) / TRD_DR, nsq = (n + TRD_CB - 1) / TRD_CB;
    dpart.ensure((size_t)nchmax * n);
    dtrW.ensure((size_t)n * TRD_NB);
    dtrDots.ensure((size_t)ndbmax * 2 * TRD_NB);
    dtrVp.ensure((size_t)nchmax * (nchmax + 1) / 2);
    dtrSq.ensure(2 * (size_t)nsq);
    dtrAx.ensure(n);
    for (int i0 = 0; i0 + 1 < n; i0 += TRD_NB) {
      const int b = std::min(TRD_NB, n - 1 - i0);
      for (int j = 0; j <= b; j++) {
        const int i = i0 + j;
        double* sq_cur = dtrSq.p + (size_t)(i & 1) * nsq;
        const double* sq_prev = dtrSq.p + (size_t)((i + 1) & 1) * nsq;
        LAUNCH(k_trd_col, dim3(blocks(n - i, TRD_CB)), dim3(256), 0, A, n, i0, j, dtrW.p, (const double*)dpart.p,
               (const double*)dtrDots.p, (const double*)dtrVp.p, sq_prev, sq_cur, dtrAx.p, tau, dd.p, de.p,
               j == b ? 1 : 0);
        if (j == b) break;
        const int m = n - i - 1;
        const double bytes = 4.0 * m * (double)(m + 1);  // the lower triangle of A22
        const int nbt = blocks(m, TRD_CH), nt = nbt * (nbt + 1) / 2;
        const int tpw = nt >= trd_pair_min ? 2 : 1;  // tiles per workgroup
        LAUNCHW(bytes, k_trd_symv, dim3(blocks(nt, tpw) + blocks(m, TRD_DR) * blocks(2 * j, 4)), dim3(256), 0,
                (const double*)A, n, i0, j, (const double*)dtrW.p, (const double*)sq_cur, (const double*)dtrAx.p,
                dpart.p, dtrDots.p, dtrVp.p, tpw);
      }
      const int r0 = i0 + b, mt = n - r0;
      double* A22 = A + (long long)r0 * n + r0;
      const double* V = A + (long long)i0 * n + r0;
      const double* Wt = dtrW.p + r0;
      // A22 -= V W^T + W V^T in one pass over A22, lower triangle (the only one read later)
      GemmJob g{mt, mt, b, 0, 1, V, Wt, A22, nullptr, n, n, n, -1.0, 1.0};
      g.A2 = Wt;
      g.B2 = V;
      g.k2 = b;
      g.lower = 1;
      std::vector<GemmJob> g1{g};
      gemm(g1);
    }
    LAUNCH(k_hh_last, dim3(1), dim3(1), 0, (const double*)A, n, dd.p, de.p, tau);
    HIP_OK(hipGetLastError());
    // 2. divide and conquer on (d, e): the merge tree by recursive halving (leaves <= 32)
    std::vector<double> hd(n), he(n);
    HIP_OK(hipMemcpyAsync(hd.data(), dd.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(he.data(), de.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    std::vector<std::vector<DcNode>> levels;  // levels[0] = root
    std::vector<DcNode> cur{DcNode{0, n, n / 2}};
    std::vector<int> ls, lm;
    while (!cur.empty()) {
      std::vector<DcNode> next, merges;
      for (auto& N : cur) {
        if (N.m <= 32) {
          ls.push_back(N.s);
          lm.push_back(N.m);
        } else {
          DcNode M{N.s, N.m, N.m / 2};
          merges.push_back(M);
          next.push_back(DcNode{N.s, M.m1, 0});
          next.push_back(DcNode{N.s + M.m1, N.m - M.m1, 0});
          // tearing: T = diag(T1', T2') + |beta| u u^T
          double b = std::fabs(he[N.s + M.m1 - 1]);
          hd[N.s + M.m1 - 1] -= b;
          hd[N.s + M.m1] -= b;
        }
      }
      if (!merges.empty()) levels.push_back(merges);
      cur = next;
    }
    DBuf<double>& dt = dkd;  // torn diagonal (dkd is free until the merges)
    dt.upload(hd, st, &stage);
    const int nleaf = (int)ls.size();
    dls.upload(ls, st, &stage);
    dlm.upload(lm, st, &stage);
    dleaf.ensure((size_t)nleaf * 32 * 32);
    dleafV.ensure((size_t)nleaf * 32 * 32);
    dleafZ.ensure((size_t)nleaf * 32 * 32);
    dleafw.ensure((size_t)nleaf * 32);
    LAUNCH(k_dc_leaves, dim3(nleaf), dim3(256), 0, (const int*)dls.p, (const int*)dlm.p, (const double*)dt.p,
           (const double*)de.p, dleaf.p, 32);
    std::vector<EigJob> jobs(nleaf);
    for (int l = 0; l < nleaf; l++)
      jobs[l] = EigJob{lm[l], dleaf.p + (size_t)l * 1024, dleafV.p + (size_t)l * 1024, dleafw.p + (size_t)l * 32,
                       dleafZ.p + (size_t)l * 1024};
    eigh(jobs, false, 60);
    dq_Q2.ensure((size_t)n * n);
    fill(Q, (long long)n * n);
    fill(dq_Q2.p, (long long)n * n);
    LAUNCH(k_dc_scatter, dim3(nleaf), dim3(256), 0, (const int*)dls.p, (const int*)dlm.p, (const double*)dleafZ.p,
           (const double*)dleafw.p, 32, Q, n, dD.p);
    // merges, deepest level first
    for (int lv = (int)levels.size() - 1; lv >= 0; lv--) {
      auto& nodes = levels[lv];
      const int nn = (int)nodes.size();
      dnodes.upload(nodes, st, &stage);
      dnk.ensure(nn);
      dnr.ensure(nn);
      drho.ensure(nn);
      int mmax = 0;
      for (auto& N : nodes) mmax = std::max(mmax, N.m);
      LAUNCH(k_dc_prep, dim3(nn), dim3(256), 0, (const DcNode*)dnodes.p, (const double*)Q, n,
             (const double*)dD.p, (const double*)de.p, dds.p, dzs.p, dperm.p, dkind.p, drot_ij.p, drot_cs.p, dnk.p,
             dnr.p, drho.p);
      std::vector<int> hk(nn), hr(nn);
      HIP_OK(hipMemcpyAsync(hk.data(), dnk.p, nn * sizeof(int), hipMemcpyDeviceToHost, st));
      HIP_OK(hipMemcpyAsync(hr.data(), dnr.p, nn * sizeof(int), hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
      int kmax = 0, rmax = 0;
      std::vector<long long> uo(nn), go(nn);
      long long ut = 0, gt = 0;
      for (int i = 0; i < nn; i++) {
        kmax = std::max(kmax, hk[i]);
        rmax = std::max(rmax, hr[i]);
        uo[i] = ut;
        go[i] = gt;
        ut += (long long)hk[i] * hk[i];
        gt += (long long)nodes[i].m * hk[i];
      }
      duoff.upload(uo, st, &stage);
      dgoff.upload(go, st, &stage);
      dU.ensure(std::max<long long>(ut, 1));
      dG.ensure(std::max<long long>(gt, 1));
      dQG.ensure(std::max<long long>(gt, 1));
      if (rmax > 0)
        LAUNCH(k_dc_rot, dim3(blocks(mmax, 256), nn), dim3(256), 0, (const DcNode*)dnodes.p, (const int*)dnr.p,
               (const int*)drot_ij.p, (const double*)drot_cs.p, Q, n);
      LAUNCH(k_dc_compact, dim3(nn), dim3(64), 0, (const DcNode*)dnodes.p, (const int*)dkind.p, (const double*)dds.p,
             (const double*)dzs.p, (const int*)dperm.p, dkd.p, dkz.p, dkcol.p);
      if (kmax > 0) {
        LAUNCH(k_dc_secular, dim3(blocks(kmax, 4), nn), dim3(256), 0, (const DcNode*)dnodes.p, (const int*)dnk.p,
               (const double*)drho.p, (const double*)dkd.p, (const double*)dkz.p, dtaus.p, dorg.p);
        LAUNCH(k_dc_zhat, dim3(blocks(kmax, 64), nn), dim3(64), 0, (const DcNode*)dnodes.p, (const int*)dnk.p,
               (const double*)drho.p, (const double*)dkd.p, (const double*)dkz.p, (const double*)dtaus.p,
               (const int*)dorg.p, dzh.p);
        LAUNCH(k_dc_u, dim3(kmax, nn), dim3(256), 0, (const DcNode*)dnodes.p, (const int*)dnk.p,
               (const long long*)duoff.p, (const double*)dkd.p, (const double*)dzh.p, (const double*)dtaus.p,
               (const int*)dorg.p, dU.p);
        LAUNCH(k_dc_gather, dim3(blocks((long long)mmax * kmax, 256), nn), dim3(256), 0, (const DcNode*)dnodes.p,
               (const int*)dnk.p, (const long long*)dgoff.p, (const int*)dkcol.p, (const double*)Q, n, dG.p);
        std::vector<GemmJob> gjobs;
        for (int i = 0; i < nn; i++)
          if (hk[i] > 0) {
            int m = nodes[i].m, K = hk[i];
            gjobs.push_back(GemmJob{m, K, K, 0, 0, dG.p + go[i], dU.p + uo[i], dQG.p + go[i], nullptr, m, K, m, 1.0, 0.0});
          }
        gemm(gjobs);
      }
      LAUNCH(k_dc_vals, dim3(blocks(nn, 64)), dim3(64), 0, (const DcNode*)dnodes.p, nn, (const int*)dnk.p,
             (const int*)dkind.p, (const double*)dds.p, (const int*)dperm.p, (const double*)dkd.p,
             (const double*)dtaus.p, (const int*)dorg.p, dval.p, dsrc.p);
      LAUNCH(k_dc_place, dim3(mmax, nn), dim3(256), 0, (const DcNode*)dnodes.p, (const long long*)dgoff.p,
             (const double*)dval.p, (const int*)dsrc.p, (const double*)dQG.p, (const double*)Q, n, dq_Q2.p,
             dD2.p);
      LAUNCH(k_dc_copyback, dim3(mmax, nn), dim3(256), 0, (const DcNode*)dnodes.p, (const double*)dq_Q2.p,
             (const double*)dD2.p, n, Q, dD.p);
      HIP_OK(hipGetLastError());
    }
    HIP_OK(hipMemcpyAsync(w, dD.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  }

  // The first k eigenvectors of the last dense_values() matrix into Z (n x k, ld n): the
  // tridiagonal eigenvectors transformed back with the Householder reflectors, nb per panel
  // (Z <- (I - V T V^T) Z per panel, last panel first).
  void dense_vectors(const double* A, int n, const double* tau, const double* Q, int k, double* Z) {
    if (k <= 0) return;
    HIP_OK(hipMemcpyAsync(Z, Q, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice, st));
    const int nb = 64;
    const int nref = n - 1;  // reflectors 0..n-2
    dbtV.ensure((size_t)n * nb);
    dbtY.ensure((size_t)nb * k);
    dbtVtV.ensure((size_t)nb * nb);
    dbtVT.ensure((size_t)n * nb);
    int np = (nref + nb - 1) / nb;
    for (int p = np - 1; p >= 0; p--) {
      const int i0 = p * nb;
      const int b = std::min(nb, nref - i0);
      const int rows = n - i0 - 1;
      LAUNCH(k_bt_panel, dim3(blocks((long long)rows * b, 256)), dim3(256), 0, A, tau, n, i0, b, dbtV.p);
      double* Zr = Z + i0 + 1;
      // V^T V and Yt = Zr^T V; T from V^T V; VT = V T; Zr -= VT Yt^T
      // two launches, so that each long-k product takes the split-K path
      std::vector<GemmJob> j1{GemmJob{b, b, rows, 1, 0, dbtV.p, dbtV.p, dbtVtV.p, nullptr, rows, rows, b, 1.0, 0.0}};
      gemm(j1);
      std::vector<GemmJob> j1b{GemmJob{k, b, rows, 1, 0, Zr, dbtV.p, dbtY.p, nullptr, n, rows, k, 1.0, 0.0}};
      gemm(j1b);
      LAUNCH(k_bt_tfactor, dim3(1), dim3(64), 0, dbtVtV.p, tau, i0, b);
      std::vector<GemmJob> j2{GemmJob{rows, b, b, 0, 0, dbtV.p, dbtVtV.p, dbtVT.p, nullptr, rows, b, rows, 1.0, 0.0}};
      gemm(j2);
      std::vector<GemmJob> j3{GemmJob{rows, k, b, 0, 1, dbtVT.p, dbtY.p, Zr, nullptr, rows, k, n, -1.0, 1.0}};
      gemm(j3);
    }
    HIP_OK(hipGetLastError());
  }

  void fill(double* p, long long n, double v = 0.0) {
    if (n <= 0) return;
    LAUNCH(k_fill, dim3(std::min<long long>(blocks(n, 256), 65535)), dim3(256), 0, p, n, v);
  }

  // ------------------------------------------------------------------------------ one batch
  // a given start must be finite (calculator.rs:482-509); the elements are checked before
  static bool finite_start(const MolIn& m) {
    if (!m.gq) return true;
    int nsh = 0;
    for (int a = 0; a < m.nat; a++) nsh += GFN2_NSHELL[m.z[a] - 1];
    bool fin = true;
    for (int k = 0; k < nsh; k++) fin = fin && std::isfinite(m.gq[k]);
    for (int k = 0; m.gdp && k < 3 * m.nat; k++) fin = fin && std::isfinite(m.gdp[k]);
    for (int k = 0; m.gqp && k < 6 * m.nat; k++) fin = fin && std::isfinite(m.gqp[k]);
    return fin;
  }

  // new positions and starts for the molecules of the prepared batch (same atoms in the same
  // order; `in` indexed as at prepare). Everything is checked before anything changes: a
  // mismatch or an invalid geometry or start fails the call and leaves the batch as it was.
  void update(const std::vector<MolIn>& in) {
    HIP_OK(hipSetDevice(device));
    if (!resident) throw Fail{AXTB_ERR_INPUT, "no prepared batch: batch_prepare first"};
    if (in.size() != prep_out.size()) throw Fail{AXTB_ERR_INPUT, "the molecule count differs from the prepared batch"};
    for (int m = 0; m < (int)mols.size(); m++) {
      const Mol& M = mols[m];
      if (map_use[m] >= (int)in.size()) throw Fail{AXTB_ERR_INPUT, "the molecule count differs from the prepared batch"};
      const MolIn& x = in[map_use[m]];
      const std::string id = "molecule " + std::to_string(map_use[m]) + ": ";
      if (x.nat != M.nat) throw Fail{AXTB_ERR_INPUT, id + "the atom count differs from the prepared batch"};
      // the molecule (numbers, charge, unpaired electrons) is fixed at prepare, as a calculator's
      if (x.charge != M.charge || x.uhf != huhf[m])
        throw Fail{AXTB_ERR_INPUT, id + "the charge or unpaired electrons differ from the prepared batch"};
      for (int a = 0; a < M.nat; a++) {
        if (x.z[a] != hz[M.at0 + a]) throw Fail{AXTB_ERR_INPUT, id + "the elements differ from the prepared batch"};
        for (int k = 0; k < 3; k++)
          if (!std::isfinite(x.xyz[3 * a + k])) throw Fail{AXTB_ERR_INPUT, id + "non-finite coordinate"};
        for (int c = 0; c < a; c++) {
          double d2 = 0;
          for (int k = 0; k < 3; k++) d2 += std::pow(x.xyz[3 * a + k] - x.xyz[3 * c + k], 2);
          if (!(d2 >= 1e-6)) throw Fail{AXTB_ERR_INPUT, id + "atoms " + std::to_string(c) + " and " + std::to_string(a) + " coincide"};
        }
      }
      if (!finite_start(x)) throw Fail{AXTB_ERR_INPUT, id + "the SCF start has non-finite values"};
    }
    for (int m = 0; m < (int)mols.size(); m++) {
      const Mol& M = mols[m];
      const MolIn& x = in[map_use[m]];
      std::copy(x.xyz, x.xyz + 3 * M.nat, hxyz.begin() + 3 * M.at0);
      hgflag[m] = x.gq != nullptr;
      for (int k = 0; k < M.nsh; k++) hgq[M.sh0 + k] = x.gq ? x.gq[k] : 0.0;
      for (int k = 0; k < 3 * M.nat; k++) hgdp[3 * M.at0 + k] = x.gq && x.gdp ? x.gdp[k] : 0.0;
      for (int k = 0; k < 6 * M.nat; k++) hgqp[6 * M.at0 + k] = x.gq && x.gqp ? x.gqp[k] : 0.0;
    }
    if (mols.empty()) return;
    a_xyz.upload(hxyz, st, &stage);
    any_guess = std::find(hgflag.begin(), hgflag.end(), 1) != hgflag.end();
    if (any_guess) {
      g_flag.upload(hgflag, st, &stage);
      g_q.upload(hgq, st, &stage);
      g_dp.upload(hgdp, st, &stage);
      g_qp.upload(hgqp, st, &stage);
    }
    HIP_OK(hipStreamSynchronize(st));
    scf_done = grad_done = false;
  }

  // validate and upload a batch (per-molecule input errors are reported in `out`)
  void prepare(const std::vector<MolIn>& in, const axtb_opts& o, std::vector<MolOut>& out) {
    HIP_OK(hipSetDevice(device));
    // validation (base/src/molecule.rs:35-64, xtb.rs:447,1159-1212)
    if (!(o.accuracy > 0.0 && std::isfinite(o.accuracy))) throw Fail{AXTB_ERR_INPUT, "accuracy must be positive"};
    if (!(o.etemp >= 0.0 && std::isfinite(o.etemp))) throw Fail{AXTB_ERR_INPUT, "electronic temperature must be >= 0 K"};
    if (!(o.damping > 0.0 && o.damping <= 1.0)) throw Fail{AXTB_ERR_INPUT, "damping must be in (0, 1]"};
    if (o.max_iter < 1) throw Fail{AXTB_ERR_INPUT, "max_iter must be >= 1"};
    if (o.mixer < AXTB_MIXER_AUTO || o.mixer > AXTB_MIXER_BROYDEN_DIIS) throw Fail{AXTB_ERR_INPUT, "unknown SCF mixer"};
    if (o.conventions != AXTB_CONV_AXTB && o.conventions != AXTB_CONV_TBLITE)
      throw Fail{AXTB_ERR_INPUT, o.conventions == AXTB_CONV_XTB || o.conventions == AXTB_CONV_XTB671
                                     ? "the GPU engine runs the axtb and tblite conventions, not the xtb programs'"
                                     : "unknown conventions"};
    if (o.precision < AXTB_PREC_FP64 || o.precision > AXTB_PREC_FP64_REFINE)
      throw Fail{AXTB_ERR_INPUT, "unknown precision"};
    if (o.solvation < AXTB_SOLV_NONE || o.solvation > AXTB_SOLV_GBSA) throw Fail{AXTB_ERR_INPUT, "unknown solvation model"};
    if (o.solvation != AXTB_SOLV_NONE) {
      if (o.solvation_state < AXTB_SOLV_GSOLV || o.solvation_state > AXTB_SOLV_REFERENCE)
        throw Fail{AXTB_ERR_INPUT, "unknown solvation reference state"};
      if (axtb_solv::find_set(o.solvent, o.solvation) < 0)
        throw Fail{AXTB_ERR_INPUT, std::string("solvation: solvent '") + (o.solvent ? o.solvent : "") +
                                       "' is not parametrized for GFN2-xTB/" +
                                       (o.solvation == AXTB_SOLV_ALPB ? "ALPB" : "GBSA")};
    }
    resident = false;
    out.assign(in.size(), MolOut());
    std::vector<int> use;  // input index of each batch molecule
    for (size_t i = 0; i < in.size(); i++) {
      const MolIn& m = in[i];
      MolOut& r = out[i];
      if (m.nat < 1) {
        r.status = AXTB_ERR_INPUT;
        r.msg = "no atoms";
        continue;
      }
      bool ok = true;
      for (int a = 0; a < m.nat && ok; a++) {
        if (m.z[a] < 1 || m.z[a] > 118) {  // not an element (base/src/molecule.rs:41-43)
          r.status = AXTB_ERR_INPUT;
          r.msg = "invalid atomic number " + std::to_string(m.z[a]);
          ok = false;
        } else if (!supported_element(m.z[a])) {  // beyond the parametrisation (gfn2.rs:37-39)
          r.status = AXTB_ERR_UNSUPPORTED_ELEMENT;
          r.msg = "element " + std::to_string(m.z[a]) + " is not supported by the GPU engine";
          r.element = m.z[a];
          ok = false;
        }
        for (int k = 0; k < 3 && ok; k++)
          if (!std::isfinite(m.xyz[3 * a + k])) {
            r.status = AXTB_ERR_INPUT;
            r.msg = "non-finite coordinate";
            ok = false;
          }
      }
      for (int a = 0; a < m.nat && ok; a++)
        for (int b = 0; b < a && ok; b++) {
          double d2 = 0;
          for (int k = 0; k < 3; k++) d2 += std::pow(m.xyz[3 * a + k] - m.xyz[3 * b + k], 2);
          if (!(d2 >= 1e-6)) {
            r.status = AXTB_ERR_INPUT;
            r.msg = "atoms " + std::to_string(b) + " and " + std::to_string(a) + " coincide";
            ok = false;
          }
        }
      if (!ok) continue;
      if (std::fmod(m.charge, 1.0) != 0.0 || !std::isfinite(m.charge)) {
        r.status = AXTB_ERR_INPUT;
        r.msg = "the charge must be an integer";
        continue;
      }
      if (!finite_start(m)) {
        r.status = AXTB_ERR_INPUT;
        r.msg = "the SCF start has non-finite values";
        continue;
      }
      use.push_back((int)i);
    }
    if (use.empty()) {
      mols.clear();
      return;
    }
    // largest molecules first: every per-molecule launch (products, eigensolvers, refinement) then
    // deals its longest jobs out first and ends on the short ones
    {
      std::vector<int> nao(in.size(), 0);
      for (int i : use)
        for (int a = 0; a < in[i].nat; a++) {
          const int iz = in[i].z[a] - 1;
          for (int s = 0; s < GFN2_NSHELL[iz]; s++) nao[i] += 2 * GFN2_ANG_SHELL[iz][s] + 1;
        }
      std::stable_sort(use.begin(), use.end(), [&](int a, int b) { return nao[a] > nao[b]; });
    }
    build(in, use, o, out);
    // the options outlive the call: the solvent's name is kept here; the mixer of the conventions'
    // default resolved (xtb.rs:516: axtb Broyden then DIIS, tblite Broyden; conventions.rs:260,348)
    solvent_name = o.solvent ? o.solvent : "";
    opts_ = o;
    opts_.solvent = solvent_name.c_str();
    if (o.mixer == AXTB_MIXER_AUTO && o.conventions == AXTB_CONV_TBLITE) opts_.mixer = AXTB_MIXER_BROYDEN;
    if (o.mixer == AXTB_MIXER_BROYDEN_DIIS) opts_.mixer = AXTB_MIXER_AUTO;
  }

  // the prepared batch on the device: setup, SCC and gradient (inputs resident in HBM)
  bool scf_done = false, grad_done = false;
  // the resident batch of batch_prepare (valid until a one-shot run re-prepares the engine) and
  // the per-input results of its validation
  bool resident = false;
  std::vector<MolOut> prep_out;
  std::vector<Attempts> att;
  // iterations and outcome of attempt k of the molecules in mask, from the iteration records
  void record_attempt(int k, int kind, const std::vector<int>& mask) {
    const int NM = nmol();
    std::vector<double> hm(8 * (size_t)NM);
    HIP_OK(hipMemcpyAsync(hm.data(), mst.p, sizeof(double) * 8 * NM, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    for (int m = 0; m < NM; m++) {
      if (!mask[m]) continue;
      att[m].n = k + 1;
      att[m].kind[k] = kind;
      att[m].iter[k] = (int)hm[8 * m + 5];
      att[m].conv[k] = hm[8 * m + 6] == 1.0;
    }
  }
  void compute() {
    scf_done = grad_done = false;
    if (mols.empty()) return;
    setup(opts_);
    // the mixer plan (xtb.rs:590-646): one attempt per mixer, each from the start; auto runs DIIS
    // for the molecules Broyden did not converge (not for model errors: a non-finite energy or an
    // eigensolver failure ends the calculation, as the reference's `?` does)
    const int NM = nmol();
    std::vector<int> all(NM, 1);
    att.assign(NM, Attempts());
    const int first = opts_.mixer == AXTB_MIXER_DIIS ? MIX_DIIS : MIX_BROYDEN;
    scf(opts_, first, &all);
    record_attempt(0, first, all);
    if (opts_.mixer == AXTB_MIXER_AUTO) {
      std::vector<int> hst(NM), again(NM, 0);
      HIP_OK(hipMemcpyAsync(hst.data(), status.p, sizeof(int) * NM, hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
      bool any = false;
      for (int m = 0; m < NM; m++)
        if (!att[m].conv[0] && (hst[m] & ~(2 | 8)) == 0) any = again[m] = 1;
      if (any) {
        scf(opts_, MIX_DIIS, &again);
        record_attempt(1, MIX_DIIS, again);
      }
    }
    scf_done = true;
    if (opts_.want_gradient) {
      gradient();
      grad_done = true;
    }
    HIP_OK(hipStreamSynchronize(st));
  }

  // the analytic gradient of the state the last compute() left (P, W's orbitals, potentials);
  // a compute() without gradient followed by this equals a compute() with it
  void gradient_of_state() {
    if (mols.empty() || grad_done) return;
    if (!scf_done) throw Fail{AXTB_ERR_INPUT, "no computed batch: batch_run first"};
    gradient();
    grad_done = true;
    HIP_OK(hipStreamSynchronize(st));
  }

  // a one-shot calculation (the engine's resident batch is replaced and no longer usable)
  void run(const std::vector<MolIn>& in, const axtb_opts& o, std::vector<MolOut>& out) {
    prepare(in, o, out);
    compute();
    if (!mols.empty()) collect(o, out);
  }
  // the resident batch: prepared once, then computed by batch_run / updated by batch_update
  void prepare_resident(const std::vector<MolIn>& in, const axtb_opts& o) {
    std::vector<MolOut> out;
    prepare(in, o, out);
    scf_done = grad_done = false;
    prep_out.swap(out);
    resident = true;
  }
  axtb_opts opts_{};
  std::string solvent_name;

  // ------------------------------------------------------------------------------ tables
  std::vector<int> hz, hsp, hmol, hish, hnsh, hnref, hngw, hsat, hsl, hsnp, hsiao, haosh, haoat;
  std::vector<double> hxyz, hrcov, hen, hrad, hra, hrz, hdk, hqk, har, hav, hchi, heta, hkcn, herad, hreeq, hd4rcov,
      hd4zeff, hd4gi, hd4r4r2, hd4refcn, hd4refq, hsal, hsco, hse0, hskcn, hsshp, hsref, hshard, hsgam3, hsmin, hszeta;
  std::vector<int> map_use;  // batch molecule -> input index
  std::vector<int> by_size;  // batch molecules by basis size, largest first
  std::vector<int> huhf;     // batch molecule -> the unpaired electrons it was prepared with
  // given starts of the batch molecules (flag per molecule; moments in the batch layout)
  std::vector<int> hgflag;
  std::vector<double> hgq, hgdp, hgqp;
  DBuf<int> g_flag;
  DBuf<double> g_q, g_dp, g_qp;
  bool any_guess = false;
  int cap = 64;
  long long tot_mat = 0, tot_pair = 0, tot_shp = 0, tot_eq = 0, tot_hist = 0;
  int tot_vec = 0;

  void build(const std::vector<MolIn>& in, const std::vector<int>& use, const axtb_opts& o,
             std::vector<MolOut>& out) {
    for (auto* v : {&hz, &hsp, &hmol, &hish, &hnsh, &hnref, &hngw, &hsat, &hsl, &hsnp, &hsiao, &haosh, &haoat}) v->clear();
    for (auto* v : {&hxyz, &hrcov, &hen, &hrad, &hra, &hrz, &hdk, &hqk, &har, &hav, &hchi, &heta, &hkcn, &herad, &hreeq,
                    &hd4rcov, &hd4zeff, &hd4gi, &hd4r4r2, &hd4refcn, &hd4refq, &hsal, &hsco, &hse0, &hskcn, &hsshp,
                    &hsref, &hshard, &hsgam3, &hsmin, &hszeta})
      v->clear();
    mols.clear();
    map_use.clear();
    huhf.clear();
    hgflag.clear();
    hgq.clear();
    hgdp.clear();
    hgqp.clear();
    cap = std::max(1, o.max_iter);  // the reference's Broyden memory: max_iter (mixing.rs:104-109)
    // the conventions (conventions.rs:244-353): axtb, or tblite 0.8.0's constants, radii, cutoffs and
    // screening
    const bool tbl = o.conventions == AXTB_CONV_TBLITE;
    tblite_ = tbl;
    auto COVALENT_RAD_D3_MCTC = [](int iz) { return AXTB_MCTC_AATOAU * EL_COVALENT_RAD_2009_AA[iz]; };
    double amin_mol = INFINITY;  // the molecule's smallest primitive exponent (basis.rs:207)
    std::map<int, int> species;
    tot_mat = tot_pair = tot_shp = tot_eq = tot_hist = 0;
    tot_vec = 0;
    int at0 = 0, sh0 = 0, ao0 = 0;
    for (int ii : use) {
      const MolIn& m = in[ii];
      Mol M{};
      M.nat = m.nat;
      M.at0 = at0;
      M.sh0 = sh0;
      M.ao0 = ao0;
      M.charge = m.charge;
      int nsh = 0, nao = 0;
      double nref = 0.0;
      for (int a = 0; a < m.nat; a++) {
        int z = m.z[a], iz = z - 1;
        if (!species.count(z)) {
          int s = (int)species.size();
          species[z] = s;
        }
        hz.push_back(z);
        hsp.push_back(species[z]);
        hmol.push_back((int)mols.size());
        hish.push_back(sh0 + nsh);
        hnsh.push_back(GFN2_NSHELL[iz]);
        for (int k = 0; k < 3; k++) hxyz.push_back(m.xyz[3 * a + k]);
        // radii in bohr: CODATA 2022 (axtb), mctc-lib's bohr (tblite; param/mod.rs:26-47,
        // element_data.rs:26,94-102)
        hrcov.push_back(tbl ? 4.0 / 3.0 * COVALENT_RAD_D3_MCTC(iz) : 4.0 / 3.0 * (EL_COVALENT_RAD_2009_AA[iz] * AATOAU_2022));
        hen.push_back(EL_PAULING_EN[iz]);
        hrad.push_back(tbl ? AXTB_MCTC_AATOAU * EL_ATOMIC_RAD_AA[iz] : EL_ATOMIC_RAD_AA[iz] * AATOAU_2022);
        hra.push_back(GFN2_REP_ALPHA[iz]);
        hrz.push_back(GFN2_REP_ZEFF[iz]);
        hdk.push_back(GFN2_P_DKERNEL[iz]);
        hqk.push_back(GFN2_P_QKERNEL[iz]);
        har.push_back(GFN2_P_RAD[iz]);
        hav.push_back(GFN2_P_VCN[iz]);
        hchi.push_back(EEQ_CHI[iz]);
        heta.push_back(EEQ_ETA[iz]);
        hkcn.push_back(EEQ_KCNCHI[iz]);
        herad.push_back(EEQ_RAD[iz]);
        hreeq.push_back(4.0 / 3.0 * (AXTB_MCTC_AATOAU * EL_COVALENT_RAD_2009_AA[iz]));
        // D4 (d4/src/model.rs:60-71, 260-340), GFN2 reference charges, axtb CN radii
        hd4rcov.push_back(tbl ? 4.0 / 3.0 * COVALENT_RAD_D3_MCTC(iz) : 4.0 / 3.0 * (EL_COVALENT_RAD_2009_AA[iz] * AATOAU_2022));
        hd4zeff.push_back(D4_EFFECTIVE_NUCLEAR_CHARGE[iz]);
        hd4gi.push_back(D4_CHEMICAL_HARDNESS[iz] * 2.0);
        hd4r4r2.push_back(std::sqrt(0.5 * (D4_R4_OVER_R2[iz] * std::sqrt((double)z))));
        int nr = D4_REFN[iz];
        hnref.push_back(nr);
        int cnc[20] = {0};
        cnc[0] = 1;
        auto icn = [](double x) { return (int)std::min<long>(19, std::max<long>(0, std::lround(x))); };
        for (int r = 0; r < nr; r++) cnc[icn(D4_REFCN[iz][r])]++;
        for (int r = 0; r < MAXREF; r++) {
          int c = r < nr ? cnc[icn(D4_REFCN[iz][r])] : 0;
          hngw.push_back(r < nr ? c * (c + 1) / 2 : 1);
          hd4refcn.push_back(r < nr ? D4_REFCOVCN[iz][r] : 0.0);
          hd4refq.push_back(r < nr ? D4_REFQ_GFN2[iz][r] : 0.0);
        }
        for (int s = 0; s < GFN2_NSHELL[iz]; s++) {
          int l = GFN2_ANG_SHELL[iz][s];
          int np = GFN2_NUMBER_OF_PRIMITIVES[iz][s];
          double al[MAXPRIM] = {0}, co[MAXPRIM] = {0};
          sto_ng(np, GFN2_PRINCIPAL_QUANTUM_NUMBER[iz][s], l, GFN2_SLATER_EXPONENT[iz][s], al, co);
          double amin = 1e300;
          for (int k = 0; k < np; k++) amin = std::min(amin, al[k]);
          amin_mol = std::min(amin_mol, amin);
          hsat.push_back(at0 + a);
          hsl.push_back(l);
          hsnp.push_back(np);
          hsiao.push_back(nao);
          for (int k = 0; k < MAXPRIM; k++) {
            hsal.push_back(al[k]);
            hsco.push_back(co[k]);
          }
          hse0.push_back(GFN2_P_SELFENERGY[iz][s]);
          hskcn.push_back(GFN2_P_KCN[iz][l]);
          hsshp.push_back(GFN2_P_SHPOLY[iz][l]);
          hsref.push_back(GFN2_REFERENCE_OCC[iz][l]);
          nref += GFN2_REFERENCE_OCC[iz][l];
          hshard.push_back(GFN2_HUBBARD_PARAMETER[iz] * GFN2_SHELL_HUBBARD[iz][l]);
          const double SHD[3] = {1.0, 0.5, 0.25};
          hsgam3.push_back(GFN2_P_HUBBARD_DERIVS[iz] * SHD[l]);
          hsmin.push_back(amin);
          hszeta.push_back(GFN2_SLATER_EXPONENT[iz][s]);
          for (int k = 0; k < 2 * l + 1; k++) {
            haosh.push_back(sh0 + nsh);
            haoat.push_back(at0 + a);
          }
          nao += 2 * l + 1;
          nsh++;
        }
      }
      // electrons (xtb.rs:1178-1212)
      double nocc = nref - m.charge;
      long nel_int = std::lround(nocc);
      double nuhf = m.uhf >= 0 ? (double)m.uhf : (double)(((nel_int % 2) + 2) % 2);
      MolOut& r = out[ii];
      if (m.uhf >= 0 && (m.uhf % 2) != (int)(((nel_int % 2) + 2) % 2)) {
        r.status = AXTB_ERR_INPUT;
        r.msg = "electron count and unpaired electrons are not compatible";
      } else if (nocc < 0.0) {
        r.status = AXTB_ERR_INPUT;
        r.msg = "the charge leaves a negative electron count";
      } else if (nuhf > nocc) {
        r.status = AXTB_ERR_INPUT;
        r.msg = "more unpaired electrons than electrons";
      } else if ((nocc + nuhf) / 2.0 > nao + 1e-9) {
        r.status = AXTB_ERR_INPUT;
        r.msg = "more electrons than the basis holds";
      }
      if (r.status) {
        // drop the molecule again
        size_t na = hz.size() - m.nat, ns = hsat.size() - nsh, no = haosh.size() - nao;
        hz.resize(na); hsp.resize(na); hmol.resize(na); hish.resize(na); hnsh.resize(na); hnref.resize(na);
        hngw.resize(na * MAXREF); hd4refcn.resize(na * MAXREF); hd4refq.resize(na * MAXREF);
        for (auto* v : {&hrcov, &hen, &hrad, &hra, &hrz, &hdk, &hqk, &har, &hav, &hchi, &heta, &hkcn, &herad, &hreeq,
                        &hd4rcov, &hd4zeff, &hd4gi, &hd4r4r2})
          v->resize(na);
        hxyz.resize(3 * na);
        hsat.resize(ns); hsl.resize(ns); hsnp.resize(ns); hsiao.resize(ns);
        hsal.resize(ns * MAXPRIM); hsco.resize(ns * MAXPRIM);
        for (auto* v : {&hse0, &hskcn, &hsshp, &hsref, &hshard, &hsgam3, &hsmin, &hszeta}) v->resize(ns);
        haosh.resize(no);
        haoat.resize(no);
        continue;
      }
      M.nsh = nsh;
      M.nao = nao;
      M.ndim = nsh + 9 * m.nat;
      M.nalpha = (nocc + nuhf) / 2.0;
      M.nbeta = (nocc - nuhf) / 2.0;
      M.kt = o.etemp > 0.0 ? o.etemp * (tbl ? KB_XTB_TBLITE : KB_HARTREE) : 0.0;
      // tblite screens the integrals' atom pairs beyond sqrt(2 intcut(acc) / alpha_min), at most
      // 40 bohr (basis.rs:139-143,247-249); axtb takes every pair
      if (tbl) {
        const double acc = std::min(std::max(o.accuracy, 1.0e-4), 1.0e3);
        const double ic = std::min(std::max(25.0 - 10.0 * std::log10(acc), 5.0), 25.0);
        const double rc = std::min(std::sqrt(2.0 * ic / amin_mol), 40.0);
        M.pair_cut2 = rc * rc;
      } else {
        M.pair_cut2 = INFINITY;
      }
      amin_mol = INFINITY;
      M.vec0 = tot_vec;
      M.mat0 = tot_mat;
      M.pair0 = tot_pair;
      M.shp0 = tot_shp;
      M.hist0 = tot_hist;
      tot_vec += M.ndim;
      tot_mat += (long long)nao * nao;
      tot_pair += (long long)m.nat * m.nat;
      tot_shp += (long long)nsh * nsh;
      tot_eq += (long long)(m.nat + 1) * (m.nat + 1);
      mols.push_back(M);
      map_use.push_back(ii);
      huhf.push_back(m.uhf);
      hgflag.push_back(m.gq != nullptr);
      for (int k = 0; k < nsh; k++) hgq.push_back(m.gq ? m.gq[k] : 0.0);
      for (int k = 0; k < 3 * m.nat; k++) hgdp.push_back(m.gq && m.gdp ? m.gdp[k] : 0.0);
      for (int k = 0; k < 6 * m.nat; k++) hgqp.push_back(m.gq && m.gqp ? m.gqp[k] : 0.0);
      at0 += m.nat;
      sh0 += nsh;
      ao0 += nao;
      out[ii].nsh = nsh;
      out[ii].nao = nao;
    }
    // the molecules by basis size, largest first: the order of the per-molecule job lists (one
    // workgroup per job: the long workgroups start first and the short ones fill the end)
    batch_gen++;
    by_size.resize(mols.size());
    for (size_t m = 0; m < mols.size(); m++) by_size[m] = (int)m;
    std::stable_sort(by_size.begin(), by_size.end(), [&](int a, int b) { return mols[a].nao > mols[b].nao; });
    // reference C6 of every species pair (d4/src/model.rs:197-258)
    int nsp = (int)species.size();
    std::vector<double> hc6((size_t)nsp * nsp * 49, 0.0);
    std::vector<std::vector<std::vector<double>>> aiw(nsp);
    for (auto& kv : species) {
      int z = kv.first, iz = z - 1, s = kv.second;
      aiw[s].assign(MAXREF, std::vector<double>(NFREQ, 0.0));
      for (int r = 0; r < D4_REFN[iz]; r++) {
        int is = D4_REFSYS[iz][r];
        if (is == 0) continue;
        double izs = D4_EFFECTIVE_NUCLEAR_CHARGE[is - 1];
        double gam = D4_CHEMICAL_HARDNESS[is - 1] * 2.0;
        double qmod = D4_REFH_GFN2[iz][r] + izs;
        double zv = qmod < 0.0 ? std::exp(3.0) : std::exp(3.0 * (1.0 - std::exp(gam * (1.0 - izs / qmod))));
        for (int k = 0; k < NFREQ; k++) {
          double a = D4_SSCALE[is - 1] * D4_SECAIW[is - 1][k] * zv;
          double v = D4_ASCALE[iz][r] * (D4_ALPHAIW[iz][r][k] - D4_HCOUNT[iz][r] * a);
          aiw[s][r][k] = v > 0.0 ? v : 0.0;
        }
      }
    }
    const double FREQ[NFREQ] = {0.000001, 0.050000, 0.100000, 0.200000, 0.300000, 0.400000, 0.500000, 0.600000,
                                0.700000, 0.800000, 0.900000, 1.000000, 1.200000, 1.400000, 1.600000, 1.800000,
                                2.000000, 2.500000, 3.000000, 4.000000, 5.000000, 7.500000, 10.00000};
    double tw[NFREQ];
    tw[0] = 0.5 * (FREQ[1] - FREQ[0]);
    for (int k = 1; k < NFREQ - 1; k++) tw[k] = 0.5 * ((FREQ[k] - FREQ[k - 1]) + (FREQ[k + 1] - FREQ[k]));
    tw[NFREQ - 1] = 0.5 * (FREQ[NFREQ - 1] - FREQ[NFREQ - 2]);
    for (auto& ki : species)
      for (auto& kj : species) {
        int si = ki.second, sj = kj.second;
        for (int ri = 0; ri < D4_REFN[ki.first - 1]; ri++)
          for (int rj = 0; rj < D4_REFN[kj.first - 1]; rj++) {
            double s = 0.0;
            for (int k = 0; k < NFREQ; k++) s += aiw[si][ri][k] * aiw[sj][rj][k] * tw[k];
            hc6[((size_t)si * nsp + sj) * 49 + ri * 7 + rj] = 3.0 / PI * s;
          }
      }
    nsp_ = nsp;
    c6ref.upload(hc6, st, &stage);
    if (mols.empty()) return;
    d_mol.upload(mols, st, &stage);
    i_z.upload(hz, st, &stage); i_sp.upload(hsp, st, &stage); i_mol.upload(hmol, st, &stage); i_ish.upload(hish, st, &stage); i_nsh.upload(hnsh, st, &stage);
    i_d4nref.upload(hnref, st, &stage); i_d4ngw.upload(hngw, st, &stage);
    s_at.upload(hsat, st, &stage); s_l.upload(hsl, st, &stage); s_np.upload(hsnp, st, &stage); s_iao.upload(hsiao, st, &stage);
    ao_sh.upload(haosh, st, &stage); ao_at.upload(haoat, st, &stage);
    a_xyz.upload(hxyz, st, &stage); a_rcov.upload(hrcov, st, &stage); a_en.upload(hen, st, &stage); a_rad.upload(hrad, st, &stage);
    a_ralpha.upload(hra, st, &stage); a_rzeff.upload(hrz, st, &stage); a_dk.upload(hdk, st, &stage); a_qk.upload(hqk, st, &stage);
    a_arad.upload(har, st, &stage); a_avcn.upload(hav, st, &stage); a_chi.upload(hchi, st, &stage); a_eta.upload(heta, st, &stage);
    a_kcn.upload(hkcn, st, &stage); a_erad.upload(herad, st, &stage); a_reeq.upload(hreeq, st, &stage); a_d4rcov.upload(hd4rcov, st, &stage);
    a_d4zeff.upload(hd4zeff, st, &stage); a_d4gi.upload(hd4gi, st, &stage); a_d4r4r2.upload(hd4r4r2, st, &stage);
    a_d4refcn.upload(hd4refcn, st, &stage); a_d4refq.upload(hd4refq, st, &stage);
    s_alpha.upload(hsal, st, &stage); s_coeff.upload(hsco, st, &stage); s_se0.upload(hse0, st, &stage); s_kcn.upload(hskcn, st, &stage);
    s_shpoly.upload(hsshp, st, &stage); s_refocc.upload(hsref, st, &stage); s_hard.upload(hshard, st, &stage); s_gam3.upload(hsgam3, st, &stage);
    s_minal.upload(hsmin, st, &stage); s_zeta.upload(hszeta, st, &stage);
    any_guess = std::find(hgflag.begin(), hgflag.end(), 1) != hgflag.end();
    if (any_guess) {
      g_flag.upload(hgflag, st, &stage);
      g_q.upload(hgq, st, &stage);
      g_dp.upload(hgdp, st, &stage);
      g_qp.upload(hgqp, st, &stage);
    }
    // primitive screening: xtb's cutoff at the accuracy (axtb), tblite's at accuracy 1 (basis.rs:135-143)
    intcut_ = tbl ? 25.0 : std::max(25.0 - 10.0 * std::log10(o.accuracy), 20.0);
    // implicit solvation: per-atom parameters of the model of every molecule (one solvent for the
    // batch; the axtb conventions' corrected model, xtb's default grid and temperature)
    solv_model = o.solvation;
    if (solv_model != AXTB_SOLV_NONE) {
      std::vector<double> hv[9];
      for (const Mol& M : mols) {
        axtb_solv::SolvModel sm;
        std::string err;
        if (!axtb_solv::build_model(M.nat, &hz[M.at0], o.solvation, o.solvent, o.solvation_state,
                                    axtb_solv::CONV_CORRECTED, axtb_solv::DEFAULT_NANG, axtb_solv::DEFAULT_TEMPERATURE,
                                    sm, err))
          throw Fail{AXTB_ERR_INPUT, "solvation: " + err};
        const std::vector<double>* src[9] = {&sm.vdwr, &sm.rho, &sm.svdw, &sm.vdwsa, &sm.trj2a,
                                             &sm.trj2b, &sm.wrp, &sm.gamsasa, &sm.hbmag};
        for (int k = 0; k < 9; k++) hv[k].insert(hv[k].end(), src[k]->begin(), src[k]->end());
        sm.vdwr.clear();  // the scalars and the grid are the same for every molecule
        solv0 = sm;
      }
      DBuf<double>* dst[9] = {&sv_vdwr, &sv_rho, &sv_svdw, &sv_vdwsa, &sv_trj2a, &sv_trj2b, &sv_wrp, &sv_gam, &sv_hbmag};
      for (int k = 0; k < 9; k++) dst[k]->upload(hv[k], st);
      std::vector<double> hg;
      for (size_t p = 0; p < solv0.weight.size(); p++) {
        for (int k = 0; k < 3; k++) hg.push_back(solv0.grid[3 * p + k]);
        hg.push_back(solv0.weight[p]);
      }
      sv_grid.upload(hg, st, &stage);
    }
  }
  // implicit solvation (ALPB / GBSA): the model (AXTB_SOLV_*), the scalars and grid of the batch's
  // solvent, per-atom parameters (build), geometry-dependent intermediates (setup)
  int solv_model = AXTB_SOLV_NONE;
  axtb_solv::SolvModel solv0;
  DBuf<double> sv_vdwr, sv_rho, sv_svdw, sv_vdwsa, sv_trj2a, sv_trj2b, sv_wrp, sv_gam, sv_hbmag, sv_grid;
  DBuf<double> sv_brad, sv_dbrad, sv_sasa, sv_wsa, sv_efix, sv_adet, sv_B, sv_work;
  DBuf<int> sv_fail;
  SolvDev solv_dev() const {
    const axtb_solv::SolvModel& S = solv0;
    return SolvDev{sv_vdwr.p, sv_rho.p, sv_svdw.p, sv_vdwsa.p, sv_trj2a.p, sv_trj2b.p, sv_wrp.p, sv_gam.p,
                   sv_hbmag.p, sv_grid.p, (int)S.weight.size(), S.alpb ? 1 : 0, S.keps, S.alpbet, S.born_scale,
                   S.lrcut * S.lrcut, S.w, S.ah0, S.ah1, S.ah3, S.gshift, sv_brad.p, sv_dbrad.p, sv_sasa.p,
                   sv_wsa.p, sv_efix.p, sv_adet.p, sv_B.p, sv_work.p};
  }
  int nsp_ = 0;
  double intcut_ = 25.0;
  bool tblite_ = false;  // tblite 0.8.0's conventions

  Batch batch() const {
    Batch b{};
    b.nmol = (int)mols.size();
    b.nat = (int)hz.size();
    b.nsh = (int)hsat.size();
    b.nao = (int)haosh.size();
    b.mol = d_mol.p;
    b.a = AtomP{i_z.p, i_sp.p, i_mol.p, i_ish.p, i_nsh.p, a_xyz.p, a_rcov.p, a_en.p, a_rad.p, a_ralpha.p,
                a_rzeff.p, a_dk.p, a_qk.p, a_arad.p, a_avcn.p, a_chi.p, a_eta.p, a_kcn.p, a_erad.p, a_reeq.p,
                a_d4rcov.p, a_d4zeff.p, a_d4gi.p, a_d4r4r2.p, a_d4refcn.p, a_d4refq.p, i_d4nref.p, i_d4ngw.p};
    b.s = ShellP{s_at.p, s_l.p, s_np.p, s_iao.p, s_alpha.p, s_coeff.p, s_se0.p, s_kcn.p, s_shpoly.p,
                 s_refocc.p, s_hard.p, s_gam3.p, s_minal.p, s_zeta.p};
    b.ao_sh = ao_sh.p;
    b.ao_at = ao_at.p;
    b.c6ref = c6ref.p;
    b.nsp = nsp_;
    b.intcut = intcut_;
    // conventions.rs:255-262 (axtb), 338-346 (tblite); D4 cutoffs disp.rs:38-47
    b.cn_cut2 = tblite_ ? 25.0 * 25.0 : 40.0 * 40.0;
    b.rep_cut2 = tblite_ ? 25.0 * 25.0 : 40.0 * 40.0;
    b.d4cn_cut2 = tblite_ ? 30.0 * 30.0 : 40.0 * 40.0;
    b.d4_rc_2b = tblite_ ? 50.0 : 72.0;
    b.d4_w_2b = tblite_ ? 0.05 : 1.0;
    b.d4_rc_3b = tblite_ ? 25.0 : 40.0;
    b.d4_w_3b = tblite_ ? 0.05 : 1.0;
    return b;
  }

  int nmol() const { return (int)mols.size(); }
  int nat() const { return (int)hz.size(); }
  int nshell() const { return (int)hsat.size(); }
  int naos() const { return (int)haosh.size(); }

  // ------------------------------------------------------------------------------ setup
  void setup(const axtb_opts& o) {
    Batch b = batch();
    const int NA = nat(), NS = nshell(), NO = naos(), NM = nmol();
    int nshmax = 0, naomax = 0, natmax = 0;
    for (auto& M : mols) {
      nshmax = std::max(nshmax, M.nsh);
      naomax = std::max(naomax, M.nao);
      natmax = std::max(natmax, M.nat);
    }
    for (auto* d : {&cn, &cneeq, &cnd4, &mrad, &dmrdcn, &qat, &qat_o, &vat, &eband, &eel, &e3, &dcn3, &erep, &dedcn,
                    &dedcn4})
      d->ensure(NA);
    for (auto* d : {&dp, &dp_o, &vdp, &g3, &grep, &grad}) d->ensure(3 * (size_t)NA);
    for (auto* d : {&qp, &qp_o, &vqp}) d->ensure(6 * (size_t)NA);
    eterm.ensure(8 * (size_t)NA);
    gw.ensure(MAXREF * (size_t)NA);
    dgw.ensure(MAXREF * (size_t)NA);
    for (auto* d : {&se, &dsedcn, &qstart, &qsh, &qsh_o, &vsh, &gq}) d->ensure(NS);
    for (auto* d : {&w, &focc}) d->ensure(NO);
    fwork.ensure(2 * (size_t)NO);
    cols.ensure(11 * (size_t)NO);
    gam.ensure(tot_shp);
    pool.ensure(11 * (size_t)tot_mat);
    {
      std::vector<int2> tl;  // the lower 32 x 32 tiles (I >= J) of the per-molecule matrices (k_fock)
      for (int m = 0; m < (int)mols.size(); m++) {
        const int nt = (mols[m].nao + 31) / 32;
        for (int I = 0; I < nt; I++)
          for (int J = 0; J <= I; J++) tl.push_back(make_int2(m, (I << 16) | J));
      }
      mat_tiles.upload(tl, st, &stage);
      n_mat_tiles = (int)tl.size();
    }
    for (auto* d : {&X, &Bm, &F, &T, &Am, &Vw, &Z, &C, &P, &Wm}) d->ensure(tot_mat);
    atmP.ensure(8 * (size_t)tot_pair);
    for (auto* d : {&eeqA, &eeqV, &eeqZ}) d->ensure(tot_eq);
    for (auto* d : {&eeqw, &eeqr, &eeqy}) d->ensure(NA + NM);
    for (auto* d : {&qin, &qout, &dq, &qlast, &dqlast}) d->ensure(tot_vec);
    fst.ensure(2 * (size_t)NM);
    mst.ensure(8 * (size_t)NM);
    for (auto* d : {&active, &status, &kocc, &itc, &eeqfail}) d->ensure(NM);
    eq0.ensure(NM);
    std::vector<long long> heq0(NM);
    long long off = 0;
    for (int m = 0; m < NM; m++) {
      heq0[m] = off;
      off += (long long)(mols[m].nat + 1) * (mols[m].nat + 1);
    }
    eq0.upload(heq0, st, &stage);
    std::vector<int> ones(NM, 1), zeros(NM, 0);
    active.upload(ones, st, &stage);
    status.upload(zeros, st, &stage);
    itc.upload(zeros, st, &stage);
    eeqfail.upload(zeros, st, &stage);
    fill(mst.p, 8LL * NM);

    // coordination numbers, levels, AES radii
    LAUNCH(k_atom_setup, dim3(blocks(NA, 4)), dim3(256), 0, b, (const int*)nullptr, cn.p, cneeq.p,
                       cnd4.p, se.p, dsedcn.p, mrad.p, dmrdcn.p);
    LAUNCH(k_gamma, dim3(blocks((long long)nshmax * nshmax, 256), NM), dim3(256), 0, b, gam.p);
    // integrals
    fill(pool.p, 11 * tot_mat);
    double ibytes = 0.0;  // the 11 matrices written once
    for (auto& M : mols) ibytes += 11.0 * 8.0 * M.nao * (double)M.nao;
    {
      // shell-pair lists by type (k_int_keys), grouped by molecule m % INT_GROUPS
      std::vector<long long> go(INT_GROUPS, 0);
      for (int m = 0; m < NM; m++)
        if (m % INT_GROUPS + 1 < INT_GROUPS) go[m % INT_GROUPS + 1] += (long long)mols[m].nsh * mols[m].nsh;
      for (int g = 1; g < INT_GROUPS; g++) go[g] += go[g - 1];
      intgoff.upload(go, st, &stage);
      intcnt.ensure(NKEYS * INT_GROUPS);
      intlist.ensure(NKEYS * (size_t)tot_shp);
      LAUNCH(k_zero_u32, dim3(blocks(NKEYS * INT_GROUPS, 256)), dim3(256), 0, (unsigned*)intcnt.p, NKEYS * INT_GROUPS);
      LAUNCH(k_int_keys, dim3(blocks((long long)nshmax * nshmax, 64), NM), dim3(64), 0, b, intcnt.p, intlist.p,
             (long long)tot_shp, (const long long*)intgoff.p);
      std::vector<int> hc(NKEYS * INT_GROUPS);
      HIP_OK(hipMemcpyAsync(hc.data(), intcnt.p, sizeof(int) * hc.size(), hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
      auto ikey = [&](int k, int nch, auto kern) {
        int mx = 0;
        for (int g = 0; g < INT_GROUPS; g++) mx = std::max(mx, hc[k * INT_GROUPS + g]);
        if (mx <= 0) return;
        ProfScope ps_(&prof, st, "k_integrals", k == 0 ? ibytes : 0.0);
        hipLaunchKernelGGL(kern, dim3(blocks(mx, 64), INT_GROUPS, nch), dim3(64), 0, st, b,
                           (const int2*)intlist.p + (long long)k * tot_shp, (const int*)intcnt.p + k * INT_GROUPS,
                           (const long long*)intgoff.p, (const double*)se.p, pool.p, (long long)tot_mat);
      };
      ikey(0, 1, k_integrals_key<0, 0>);
      ikey(1, 1, k_integrals_key<0, 1>);
      ikey(2, 1, k_integrals_key<0, 2>);
      ikey(3, 1, k_integrals_key<1, 0>);
      ikey(4, 1, k_integrals_key<1, 1>);
      ikey(5, 3, k_integrals_key<1, 2>);
      ikey(6, 5, k_integrals_key<2, 0>);
      ikey(7, 5, k_integrals_key<2, 1>);
      ikey(8, 5, k_integrals_key<2, 2>);
    }
    // repulsion
    LAUNCH(k_repulsion, dim3(blocks(NA, 4)), dim3(256), 0, b, (const int*)nullptr, erep.p, grep.p);
    // implicit solvation: Born radii, surfaces, aDet, Born matrices (solvation lib.rs:555-613)
    if (solv_model != AXTB_SOLV_NONE) {
      for (auto* d : {&sv_brad, &sv_dbrad, &sv_sasa, &sv_efix, &sv_work}) d->ensure(NA);
      sv_wsa.ensure((size_t)NA * solv0.weight.size());
      sv_adet.ensure(NM);
      sv_B.ensure(tot_pair);
      sv_fail.upload(zeros, st, &stage);
      const SolvDev S = solv_dev();
      LAUNCH(k_solv_born, dim3(blocks(NA, 4)), dim3(256), 0, b, S);
      LAUNCH(k_solv_sasa, dim3(NA), dim3(64), 0, b, S, sv_fail.p);
      if (S.alpb) LAUNCH(k_solv_adet, dim3(NM), dim3(256), 0, b, S);
      LAUNCH(k_solv_bmat, dim3(blocks((long long)natmax * natmax, 256), NM), dim3(256), 0, b, S);
    }
    // D4 three-body at q = 0 (weights with dgw/dCN)
    fill(fwork.p, 0);
    DBuf<double>& zq = eeqy;  // zero charges
    fill(zq.p, NA);
    LAUNCH(k_d4_weights, dim3(blocks(NA, 128)), dim3(128), 0, b, (const int*)nullptr, cnd4.p, zq.p,
                       gw.p, dgw.p, 1);
    LAUNCH(k_atm_pairs, dim3(blocks((long long)natmax * natmax, 256), NM), dim3(256), 0, b, gw.p,
                       dgw.p, atmP.p);
    if (o.dispersion) {
      atmGD.ensure(2 * (size_t)tot_pair);
      std::vector<int4> tl;  // the 16 x 16 tiles (I <= J) of every molecule's atom pairs
      for (int m = 0; m < NM; m++) {
        const int nt = (mols[m].nat + ATM_T - 1) / ATM_T;
        for (int I = 0; I < nt; I++)
          for (int J = I; J < nt; J++) tl.push_back(make_int4(m, I, J, 0));
      }
      atm_tiles.upload(tl, st, &stage);
      LAUNCH(k_atm_pairsum_tiled, dim3((unsigned)tl.size()), dim3(256), 0, b, (const int4*)atm_tiles.p,
             (const double*)atmP.p, atmGD.p);
      LAUNCH(k_atm, dim3(blocks(NA, 4)), dim3(256), 0, b, (const double*)atmP.p, (const double*)atmGD.p, e3.p, g3.p,
             dcn3.p);
    } else {
      fill(e3.p, NA);
      fill(g3.p, 3LL * NA);
      fill(dcn3.p, NA);
    }
    // start: EEQ charges (or neutral atoms); systems that fit in LDS by LU, larger ones from the
    // eigendecomposition
    if (!max_lds) {
      HIP_OK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
      if (max_lds <= 0) max_lds = 64 * 1024;
    }
    auto fits = [&](int n) { return ((size_t)n * n + n) * sizeof(double) <= (size_t)max_lds; };
    auto fits_packed = [&](int n) { return ((size_t)n * (n + 1) / 2 + n) * sizeof(double) <= (size_t)max_lds; };
    fill(qstart.p, NS);
    if (o.guess == 0 && !tblite_) {  // tblite starts from zero charges (conventions.rs:335)
      LAUNCH(k_eeq_system, dim3(blocks((long long)(natmax + 1) * (natmax + 1), 256), NM), dim3(256), 0,
                         b, eq0.p, cneeq.p, eeqA.p, eeqr.p);
      std::vector<EigJob> jobs;
      std::vector<int> lu, ev;
      int nlu = 0;
      for (int m = 0; m < NM; m++) {
        int n1 = mols[m].nat + 1;
        if (fits(n1)) {
          lu.push_back(m);
          nlu = std::max(nlu, n1);
          continue;
        }
        if (dense(n1) && eeq_cholesky(m, heq0[m])) continue;
        ev.push_back(m);
        EigJob j{n1, eeqA.p + heq0[m], eeqV.p + heq0[m], eeqw.p + mols[m].at0 + m, eeqZ.p + heq0[m]};
        if (dense(n1))
          dense_eigh(j.A, n1, j.w, j.V, j.Z);
        else
          jobs.push_back(j);
      }
      if (!lu.empty()) {
        mlu.upload(lu, st, &stage);
        const size_t shm = ((size_t)nlu * nlu + nlu) * sizeof(double);
        HIP_OK(hipFuncSetAttribute((const void*)k_eeq_lu, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        LAUNCH(k_eeq_lu, dim3(lu.size()), dim3(512), shm, b, (const int*)mlu.p, eq0.p, (const double*)eeqA.p,
               (const double*)eeqr.p, qstart.p, eeqfail.p);
      }
      if (!ev.empty()) {
        eigh(jobs);
        mev.upload(ev, st, &stage);
        LAUNCH(k_eeq_solve, dim3(ev.size()), dim3(256), 0, b, (const int*)mev.p, eq0.p, eeqZ.p, eeqw.p, eeqr.p,
               eeqy.p, qstart.p, eeqfail.p);
      }
    }
    // S-orthonormal start basis X: L^-T from a Cholesky factor in LDS where S fits (its lower
    // triangle packed where the square does not) or blocked on the matrix cores for dense-solver
    // molecules, else the Loewdin orthogonaliser U s^-1/2 U^T from the eigenpairs of S
    {
      std::vector<EigJob> jobs, cj, cjp;
      std::vector<int> lw;
      int nch = 0, nchp = 0;
      for (int m = 0; m < NM; m++) {
        long long o0 = mols[m].mat0;
        const int n = mols[m].nao;
        if (!dense(n) && fits(n)) {
          cj.push_back(EigJob{n, pool.p + o0, nullptr, nullptr, X.p + o0});
          nch = std::max(nch, n);
          continue;
        }
        if (!dense(n) && fits_packed(n)) {
          cjp.push_back(EigJob{n, pool.p + o0, nullptr, nullptr, X.p + o0});
          nchp = std::max(nchp, n);
          continue;
        }
        if (dense(n)) {
          if (!dense_chol_basis(pool.p + o0, n, Am.p + o0, Vw.p + o0, T.p + o0, X.p + o0))
            throw Fail{AXTB_ERR_NUMERICAL, "overlap matrix is not positive definite"};
          continue;
        }
        lw.push_back(m);
        HIP_OK(hipMemcpyAsync(Am.p + o0, pool.p + o0, sizeof(double) * n * n, hipMemcpyDeviceToDevice, st));
        jobs.push_back(EigJob{n, Am.p + o0, Vw.p + o0, w.p + mols[m].ao0, Z.p + o0});
      }
      if (!cj.empty()) {
        cjobs.upload(cj, st, &stage);
        cfail.ensure(cj.size());
        const size_t shm = ((size_t)nch * nch + nch) * sizeof(double);
        HIP_OK(hipFuncSetAttribute((const void*)k_chol_basis, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        LAUNCH(k_chol_basis, dim3(cj.size()), dim3(512), shm, (const EigJob*)cjobs.p, cfail.p);
        std::vector<int> hf(cj.size());
        HIP_OK(hipMemcpyAsync(hf.data(), cfail.p, sizeof(int) * cj.size(), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        for (int f : hf)
          if (f) throw Fail{AXTB_ERR_NUMERICAL, "overlap matrix is not positive definite"};
      }
      if (!cjp.empty()) {
        cjobs.upload(cjp, st, &stage);
        cfail.ensure(cjp.size());
        const size_t shm = ((size_t)nchp * (nchp + 1) / 2 + nchp) * sizeof(double);
        HIP_OK(hipFuncSetAttribute((const void*)k_chol_basis_packed, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)shm));
        LAUNCH(k_chol_basis_packed, dim3(cjp.size()), dim3(512), shm, (const EigJob*)cjobs.p, cfail.p);
        std::vector<int> hf(cjp.size());
        HIP_OK(hipMemcpyAsync(hf.data(), cfail.p, sizeof(int) * cjp.size(), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        for (int f : hf)
          if (f) throw Fail{AXTB_ERR_NUMERICAL, "overlap matrix is not positive definite"};
      }
      if (!lw.empty()) {
        eigh(jobs);
        std::vector<double> hw(NO);
        HIP_OK(hipMemcpyAsync(hw.data(), w.p, sizeof(double) * NO, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        std::vector<double> hs12(NO, 0.0);
        for (int m : lw)
          for (int i = 0; i < mols[m].nao; i++) {
            double sv = hw[mols[m].ao0 + i];
            if (!(sv > 0.0)) throw Fail{AXTB_ERR_NUMERICAL, "overlap matrix is not positive definite"};
            hs12[mols[m].ao0 + i] = 1.0 / std::sqrt(sv);
          }
        s12_.upload(hs12, st, &stage);
        std::vector<GemmJob> gjobs;
        for (int m : lw) {
          long long o0 = mols[m].mat0;
          int n = mols[m].nao;
          gjobs.push_back(GemmJob{n, n, n, 0, 1, Z.p + o0, Z.p + o0, X.p + o0, s12_.p + mols[m].ao0, n, n, n, 1.0, 0.0});
        }
        gemm(gjobs);
      }
    }
  }
  // X = L^-T for the n x n overlap S (n > 200): blocked right-looking Cholesky S = L L^T in W
  // (panels of 64: k_chol_diag factors and inverts the diagonal block, L21 = A21 L11^-T in place,
  // A22 -= L21 L21^T on its lower tiles), then M = L^-1 by recursive halving over the 64-row blocks
  // (M21 = -M22 (L21 M11), all nodes of a level in one batch; M, zero above the diagonal, in Mb),
  // X = M^T. Tt is n x n work. Returns false if a pivot is not positive.
  bool dense_chol_basis(const double* S, int n, double* W, double* Mb, double* Tt, double* X) {
    HIP_OK(hipMemcpyAsync(W, S, sizeof(double) * n * (size_t)n, hipMemcpyDeviceToDevice, st));
    fill(Mb, (long long)n * n);
    cfail.ensure(1);
    LAUNCH(k_zero_u32, dim3(1), dim3(64), 0, (unsigned*)cfail.p, 1);
    const long long ln = n;
    for (int k0 = 0; k0 < n; k0 += 64) {
      const int b = std::min(64, n - k0), rest = n - k0 - b;
      LAUNCH(k_chol_diag, dim3(1), dim3(256), 0, W, n, k0, b, Mb, cfail.p);
      if (rest == 0) break;
      double* A21 = W + (k0 + b) + k0 * ln;
      // L21 = A21 (L11^-1)^T, in place: each output tile reads its own rows, all b columns fit one tile
      std::vector<GemmJob> gp{GemmJob{rest, b, b, 0, 1, A21, Mb + k0 + k0 * ln, A21, nullptr, ln, ln, ln, 1.0, 0.0}};
      gemm(gp);
      GemmJob t{rest, rest, b, 0, 1, A21, A21, W + (k0 + b) + (k0 + b) * ln, nullptr, ln, ln, ln, -1.0, 1.0};
      t.lower = 1;
      std::vector<GemmJob> gt{t};
      gemm(gt);
    }
    // the inverse: nodes [s, e) of 64-row blocks, split at the middle; leaves (one block) are done
    const int nb = (n + 63) / 64;
    std::vector<std::vector<std::pair<int, int>>> levels;
    std::vector<std::pair<int, int>> cur{{0, nb}};
    while (!cur.empty()) {
      std::vector<std::pair<int, int>> inner, next;
      for (auto& nd : cur)
        if (nd.second - nd.first > 1) {
          inner.push_back(nd);
          const int mid = (nd.first + nd.second) / 2;
          next.push_back({nd.first, mid});
          next.push_back({mid, nd.second});
        }
      if (!inner.empty()) levels.push_back(inner);
      cur.swap(next);
    }
    for (int lv = (int)levels.size() - 1; lv >= 0; lv--) {
      std::vector<GemmJob> j1, j2;
      for (auto& nd : levels[lv]) {
        const int mid = (nd.first + nd.second) / 2;
        const int r1 = nd.first * 64, r2 = mid * 64, r3 = std::min(nd.second * 64, n);
        const int h1 = r2 - r1, h2 = r3 - r2;
        // T21 = L21 M11; M21 = -M22 T21
        j1.push_back(GemmJob{h2, h1, h1, 0, 0, W + r2 + r1 * ln, Mb + r1 + r1 * ln, Tt + r2 + r1 * ln, nullptr, ln, ln,
                             ln, 1.0, 0.0});
        j2.push_back(GemmJob{h2, h1, h2, 0, 0, Mb + r2 + r2 * ln, Tt + r2 + r1 * ln, Mb + r2 + r1 * ln, nullptr, ln, ln,
                             ln, -1.0, 0.0});
      }
      gemm(j1);
      gemm(j2);
    }
    LAUNCH(k_lower_to_upper_t, dim3(blocks(n, 32), blocks(n, 32)), dim3(256), 0, (const double*)Mb, X, n);
    int hf = 0;
    HIP_OK(hipMemcpyAsync(&hf, cfail.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    return hf == 0;
  }
  // Large EEQ systems (N + 1 > 200): the N x N block A of [A 1; 1^T 0] by the blocked Cholesky
  // (A^-1 = X X^T, X = L^-T), two solves and the Schur complement of the constraint (k_eeq_schur).
  // Work in the matrix pools at the molecule's offset (free during setup; N <= n_ao). Returns
  // false, leaving nothing changed, if A is not positive definite (the eigensolver path follows).
  bool eeq_cholesky(int m, long long e0) {
    const int N = mols[m].nat, n1 = N + 1;
    const long long o0 = mols[m].mat0;
    HIP_OK(hipMemcpy2DAsync(T.p + o0, sizeof(double) * N, eeqA.p + e0, sizeof(double) * n1, sizeof(double) * N, N,
                            hipMemcpyDeviceToDevice, st));
    if (!dense_chol_basis(T.p + o0, N, Am.p + o0, Vw.p + o0, C.p + o0, X.p + o0)) return false;
    Batch b = batch();
    double* R = Z.p + o0;      // [r | 1], N x 2
    double* Y = F.p + o0;      // X^T R
    double* Xs = Wm.p + o0;    // X X^T R
    LAUNCH(k_eeq_rhs, dim3(blocks(N, 256)), dim3(256), 0, b, m, (const double*)eeqr.p, R);
    std::vector<GemmJob> g1{GemmJob{N, 2, N, 1, 0, X.p + o0, R, Y, nullptr, N, N, N, 1.0, 0.0}};
    gemm(g1);
    std::vector<GemmJob> g2{GemmJob{N, 2, N, 0, 0, X.p + o0, Y, Xs, nullptr, N, N, N, 1.0, 0.0}};
    gemm(g2);
    LAUNCH(k_eeq_schur, dim3(1), dim3(256), 0, b, m, (const double*)Xs, (const double*)eeqr.p, qstart.p, eeqfail.p);
    return true;
  }
  DBuf<int> mlu, mev, cfail;
  DBuf<EigJob> cjobs;
  DBuf<double> s12_, mtau, oa_emax, oa_delta;
  // FP64 refinement steps of the eigenvectors U (job.Z) of job.A for batched molecules; the
  // eigenvalues go to job.w. Scratch: C (A U), F (S), Vw (G), Wm (I + E). Steps alternate U between
  // Z and T. Without cls: round 0 is oa_steps steps (rounded up to a pair), round 1 a pair, each
  // accepted at max |e_ij| <= 1e-7 (the error left is about its square). With cls (the split of
  // the orbitals into occupied / empty, k_oa_step), refined from U = I: the first step is k_oa_step
  // alone (S = A, G = I, I + E is U), the later rounds are single steps, accepted at split_accept,
  // for at most split_steps steps.
  // No host read and no job list per iteration: refine_prepare uploads the step lists of every
  // batched molecule of the SCF once (both buffer parities), each job running while rd_run[m] != 0;
  // refine_run launches every job of every step (the others skip by flag), and k_refine_round stops
  // the jobs at the end of each round (accepted: in_t = 1 when the vectors ended in T, 0 in Z; after
  // the last round: fail = 1, for a solve by the FP64 Jacobi).
  DBuf<GemmJob> rd_g;
  DBuf<EigJob> rd_e;
  DBuf<int> rd_mol, rd_run;
  DBuf<unsigned long long> rd_cnt;  // [full, split] x [jobs refined, steps, failures]
  std::vector<GemmJob> rd_L[2][4];  // per parity: A U, S and G, FP32 updates, FP64 updates
  size_t rd_goff[2][4] = {};
  int rd_ne = 0;
  void refine_prepare(const std::vector<EigJob>& ea, const std::vector<int>& mol) {
    const int NE = (int)ea.size();
    rd_ne = NE;
    for (auto& par : rd_L)
      for (auto& l : par) l.clear();
    if (!NE) return;
    for (const EigJob& j : ea)
      if (j.n > 200) throw DeviceError("refine: batched refinement called for n > 200");
    oa_delta.ensure(NE);
    oa_emax.ensure(NE);
    rd_run.ensure(nmol());
    if (!rd_cnt.p) {
      rd_cnt.ensure(6);
      HIP_OK(hipMemsetAsync(rd_cnt.p, 0, 6 * sizeof(unsigned long long), st));
    }
    std::vector<EigJob> O[2], Oid;
    for (int q = 0; q < NE; q++) {
      const EigJob& j = ea[q];
      const long long o0 = j.Z - Z.p;
      const int n = j.n;
      const int* rp = rd_run.p + mol[q];
      for (int par = 0; par < 2; par++) {
        double* U = (par == 0 ? Z.p : T.p) + o0;
        double* out = (par == 0 ? T.p : Z.p) + o0;
        GemmJob au{n, n, n, 0, 0, j.A, U, C.p + o0, nullptr, n, n, n, 1.0, 0.0};
        GemmJob gs{n, n, n, 1, 0, U, C.p + o0, F.p + o0, nullptr, n, n, n, 1.0, 0.0};
        GemmJob gg{n, n, n, 1, 0, U, U, Vw.p + o0, nullptr, n, n, n, 1.0, 0.0};
        GemmJob gu{n, n, n, 0, 0, U, Wm.p + o0, out, nullptr, n, n, n, 1.0, 0.0};
        gs.lower = gg.lower = sym_lower(n, false);
        gu.f32upd = (upd32 && prec == AXTB_PREC_MIXED && n > 128) ? 1 : 0;  // small products stay FP64
        au.run = gs.run = gg.run = gu.run = rp;
        rd_L[par][0].push_back(au);
        rd_L[par][1].push_back(gs);
        rd_L[par][1].push_back(gg);
        rd_L[par][gu.f32upd ? 2 : 3].push_back(gu);
        EigJob e{n, F.p + o0, Vw.p + o0, j.w, Wm.p + o0};
        e.run = rp;
        O[par].push_back(e);
      }
      EigJob e{n, j.A, nullptr, j.w, T.p + o0};
      e.run = rp;
      Oid.push_back(e);
    }
    std::vector<GemmJob> G;
    for (int par = 0; par < 2; par++)
      for (int k = 0; k < 4; k++) {
        rd_goff[par][k] = G.size();
        G.insert(G.end(), rd_L[par][k].begin(), rd_L[par][k].end());
      }
    std::vector<EigJob> E;
    E.insert(E.end(), O[0].begin(), O[0].end());
    E.insert(E.end(), O[1].begin(), O[1].end());
    E.insert(E.end(), Oid.begin(), Oid.end());
    gmeta.clear();
    rd_g.upload(G, st, &stage);
    rd_e.upload(E, st, &stage);
    rd_mol.upload(mol, st, &stage);
  }
  // One refinement over the prepared lists: the jobs of the molecules with src[m] != 0 (their
  // in_t and fail cleared by the caller); cls as above.
  void refine_run(const int* src, int* in_t, int* fail, const int* cls) {
    const int NE = rd_ne;
    if (!NE) return;
    LAUNCH(k_set_run, dim3(blocks(NE, 256)), dim3(256), 0, NE, (const int*)rd_mol.p, src, rd_run.p);
    const int kind = cls ? 1 : 0;
    in_refine = true;
    int gstep = 0;
    // the rounds: split 2, 1, 1, ... steps; full 2 and 2
    const int nrounds = cls ? std::max(2, split_steps - 1) : 2;
    for (int round = 0; round < nrounds; round++) {
      const int nsteps = round == 0 ? std::max(2, oa_steps + (oa_steps & 1)) : (cls ? 1 : 2);
      const double accept = cls ? split_accept : 1e-7;
      int nst = 0;
      for (int step = 0; step < nsteps; step++, gstep++) {
        const bool ident = cls && gstep == 0;
        const int par = gstep % 2;
        if (ident) {
          LAUNCH(k_oa_step, dim3(NE), dim3(256), 0, (const EigJob*)rd_e.p + 2 * NE, oa_emax.p, oa_delta.p,
                 (const int*)nullptr, 1, 1, cls, (const double*)w.p, oa_ovmix);
          continue;
        }
        gemm(rd_L[par][0], rd_g.p + rd_goff[par][0]);
        gemm(rd_L[par][1], rd_g.p + rd_goff[par][1]);
        LAUNCH(k_oa_step, dim3(NE), dim3(256), 0, (const EigJob*)rd_e.p + par * NE, oa_emax.p, oa_delta.p,
               (const int*)nullptr, gstep == 0 ? 1 : 0, 0, cls, (const double*)w.p, oa_ovmix);
        if (!rd_L[par][2].empty()) gemm(rd_L[par][2], rd_g.p + rd_goff[par][2]);
        if (!rd_L[par][3].empty()) gemm(rd_L[par][3], rd_g.p + rd_goff[par][3]);
        nst++;
      }
      LAUNCH(k_refine_round, dim3(blocks(NE, 256)), dim3(256), 0, NE, (const int*)rd_mol.p, rd_run.p,
             (const double*)oa_emax.p, accept, gstep % 2, round == nrounds - 1 ? 1 : 0, nst, round == 0 ? 1 : 0, in_t,
             fail, rd_cnt.p + 3 * kind);
    }
    in_refine = false;
  }

  // The same refinement for one large matrix (the dense-solver molecules): the GEMMs on 128-tile
  // kernels, the elementwise step over many workgroups. The refined vectors end in Z; returns the
  // last max |e_ij|.
  DBuf<double> oab_part, oab_emax, oab_delta;
  double refine_big(const EigJob& j, const int* cls) {
    const long long o0 = j.Z - Z.p;
    const int n = j.n;
    const int nb = 1024;
    oab_part.ensure(3 * nb);
    oab_emax.ensure(nb);
    oab_delta.ensure(1);
    const EigJob oj{n, F.p + o0, Vw.p + o0, j.w, Wm.p + o0};
    double emax = 0.0;
    int step = 0;
    while (step < dense_steps) {
      double* U = (step % 2 == 0 ? Z.p : T.p) + o0;
      double* out = (step % 2 == 0 ? T.p : Z.p) + o0;
      std::vector<GemmJob> gau{GemmJob{n, n, n, 0, 0, j.A, U, C.p + o0, nullptr, n, n, n, 1.0, 0.0}};
      gemm(gau);
      // S = U^T (A U) and G = U^T U are symmetric: their lower tiles, then mirrored
      std::vector<GemmJob> gsg{GemmJob{n, n, n, 1, 0, U, C.p + o0, F.p + o0, nullptr, n, n, n, 1.0, 0.0},
                               GemmJob{n, n, n, 1, 0, U, U, Vw.p + o0, nullptr, n, n, n, 1.0, 0.0}};
      gsg[0].lower = gsg[1].lower = 1;
      gemm(gsg);
      const int nbm = blocks(n, 32);
      LAUNCH(k_mirror_lower, dim3(nbm * (nbm + 1) / 2), dim3(256), 0, F.p + o0, n, (long long)n);
      LAUNCH(k_mirror_lower, dim3(nbm * (nbm + 1) / 2), dim3(256), 0, Vw.p + o0, n, (long long)n);
      LAUNCH(k_oa_big_norms, dim3(nb), dim3(256), 0, oj, oab_part.p);
      LAUNCH(k_oa_big_e, dim3(nb), dim3(256), 0, oj, (const double*)oab_part.p, nb, oab_delta.p, step == 0 ? 1 : 0,
             oab_emax.p, cls, dense_ovmix);
      std::vector<GemmJob> gu{GemmJob{n, n, n, 0, 0, U, Wm.p + o0, out, nullptr, n, n, n, 1.0, 0.0}};
      gu[0].f32upd = (upd32 && prec == AXTB_PREC_MIXED) ? 1 : 0;
      gemm(gu);
      step++;
      std::vector<double> h(nb);
      HIP_OK(hipMemcpyAsync(h.data(), oab_emax.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
      emax = 0.0;
      for (double v : h) emax = std::max(emax, v);
      if (emax <= dense_accept || emax >= 1.0) break;
    }
    if (step % 2 == 1)  // odd number of steps: the result is in T
      HIP_OK(hipMemcpyAsync(j.Z, T.p + o0, sizeof(double) * n * n, hipMemcpyDeviceToDevice, st));
    prof.count("refine_steps_dense", 1, (double)step);
    return emax;
  }
  // symmetric products: small ones (n <= 128) on the lower 16 x 16 blocks of k_gemm_small, which
  // writes the mirror itself; dense-solver molecules' on their lower 128-tiles, mirrored here
  static int sym_lower(int n, bool dense_mol) { return (dense_mol || n <= 128) ? 1 : 0; }
  void mirror_dense(const std::vector<GemmJob>& jobs) {
    for (const GemmJob& j : jobs)
      if (j.lower && j.m > 128) {
        const int nbm = blocks(j.m, 32);
        LAUNCH(k_mirror_lower, dim3(nbm * (nbm + 1) / 2), dim3(256), 0, j.C, j.m, j.ldc);
      }
  }
  DBuf<double> atmGD;  // ATM pair sums (2 x the atom-pair buffer)
  DBuf<int4> atm_tiles;
  DBuf<int> dcls, dnfrac;
  std::vector<int> dense_pfp;  // per molecule: its last eigenproblem was refined (W = P F P / 2)
  int dense_steps = 4;         // mixed, dense: refinement steps at most per SCC iteration
  // mixed, dense: a refinement is accepted once its step's largest correction max |e_ij| (over the
  // occupied-empty pairs and the orthogonality terms) is <= dense_accept; the error it leaves is of
  // second order in that correction (AXTB_HIP_DENSE_ACCEPT)
  double dense_accept = 1e-3;
  // mixed, dense: largest first-order mixing of an occupied-empty pair the refinement takes on; a
  // larger one ends it (a full solve follows)
  double dense_ovmix = 0.05;
  int dense_full = 1;  // mixed: SCC iterations solved in full before refinement (AXTB_HIP_DENSE_FULL)
  // mixed: refine the previous iteration's vectors of dense-solver molecules instead of a full
  // solve (AXTB_HIP_DENSE_REFINE=0 turns it off). Only the occupied-empty pairs are resolved: P
  // depends only on the occupied subspace, and W is formed as P F P / 2 (exact when every
  // occupation is 0 or 2, the condition for refining); the orbital energies are the Rayleigh
  // quotients of the refined vectors
  bool dense_refine = true;
  bool gemm_small = true;      // products with m, n <= 128 on the one-workgroup-per-job kernel (AXTB_HIP_GEMM_SMALL)
  // mixed: the refinement updates U (I + E) as U + U E with U E on the FP32 matrix cores (AXTB_HIP_UPD32)
  bool upd32 = true;
  // large plain products on rocBLAS (AXTB_HIP_ROCBLAS, smallest m, n and k: AXTB_HIP_ROCBLAS_MIN)
  bool use_rocblas = true;
  int rocblas_min = 512;
  rocblas_handle blas = nullptr;
  bool in_refine = false;      // the refinement's batches always take the 128-tile kernel (measured faster)
  int oa_steps = 2;  // refinement steps per SCC iteration (AXTB_HIP_OA_STEPS)
  bool jacobi_wide = true;  // 1024-thread Jacobi workgroups when one fills a CU (AXTB_HIP_JACOBI_WIDE)
  // mixed, batched molecules: SCC iterations after the first refine the previous orbitals for the
  // occupied / empty split when the occupations are all 0 or 2 (AXTB_HIP_SPLIT_REFINE)
  bool split_refine = true;
  double oa_ovmix = 0.05;  // largest first-order occupied / empty mixing a split step corrects
  double split_accept = 1e-9;  // max |e_ij| of the last split step
  int split_steps = 6;  // steps of the split refinement before a full solve (AXTB_HIP_SPLIT_STEPS)
  int split_from = 3;   // first SCC iteration that may use it (AXTB_HIP_SPLIT_FROM)
  // the split refinement starts once the previous iteration's residual is below this: the coupling
  // inside the blocks (the change of F since the last full solve) sets its convergence rate
  double split_residual = 5e-2;
  DBuf<int> dmlist, ocls, onfrac;
  DBuf<BlockCopy> dbcopy;
  DBuf<int2> mat_tiles, h0list;
  DBuf<int> h0cnt;
  DBuf<long long> h0goff, intgoff;
  DBuf<int2> intlist;
  DBuf<int> intcnt;
  int n_mat_tiles = 0;

  // ------------------------------------------------------------------------------ SCC loop
  std::vector<int> hactive;
  std::vector<int> hsplit;          // batched molecule -> may refine for the split in the next iteration
  EighPlan ep64, ep32;              // the fixed eigensolver lists of an SCF without dense molecules
  std::vector<GemmJob> sl_[4];      // the fixed product lists (host copies of sjobs)
  size_t sloff_[4] = {0, 0, 0, 0};
  DBuf<int> dlist;                  // the batched molecules of the fixed lists
  std::vector<long long> list_sig;  // what the fixed lists on the device were built for
  long long batch_gen = 0;          // counts the batches built
  int* pin_act = nullptr;           // pinned: the active mask after iterations of either parity
  size_t pin_cap = 0;
  DBuf<int> splitok, rflags, rg_mol, dpfp;
  DBuf<GemmJob> sjobs;              // the per-iteration products of the batched molecules
  DBuf<EigJob> rg_jobs;
  int iters_ = 0;

  // One SCF attempt of the molecules in `mask` (all when null) with the vector mixer `mixer`
  // (MIX_BROYDEN or MIX_DIIS), each from its start (xtb.rs:590-646); the other molecules keep
  // their state.
  void scf(const axtb_opts& o, int mixer = MIX_BROYDEN, const std::vector<int>* mask = nullptr) {
    Batch b = batch();
    const int NA = nat(), NS = nshell(), NO = naos(), NM = nmol();
    int naomax = 0;
    for (auto& M : mols) naomax = std::max(naomax, M.nao);
    const long long stride = tot_mat;
    std::vector<int> in_mask = mask ? *mask : std::vector<int>(NM, 1);
    dmask.upload(in_mask, st, &stage);
    if (mixer == MIX_DIIS) {
      for (auto* d : {&dhq, &dhr, &dbas}) d->ensure((size_t)DIIS_SIZE * tot_vec);
      for (auto* d : {&dbq, &dbr}) d->ensure(tot_vec);
      ddst.ensure(2 * (size_t)NM);
      dist.ensure(4 * (size_t)NM);
    }
    LAUNCH(k_attempt_start, dim3(blocks(NM, 256)), dim3(256), 0, (const int*)dmask.p, NM, active.p, itc.p, status.p,
           mst.p, mixer == MIX_DIIS ? ddst.p : (double*)nullptr, dist.p, o.damping);
    hist_slots = 0;
    if (solv_model != AXTB_SOLV_NONE)
      LAUNCH(k_solv_flag, dim3(blocks(NM, 256)), dim3(256), 0, (const int*)sv_fail.p, NM, active.p, status.p);
    // start moments -> q_in (dipoles and quadrupoles zero); a given start replaces them
    fill(dp.p, 3LL * NA);
    fill(qp.p, 6LL * NA);
    if (any_guess)
      LAUNCH(k_start_guess, dim3(blocks(NA, 128)), dim3(128), 0, b, (const int*)g_flag.p, (const double*)g_q.p,
             (const double*)g_dp.p, (const double*)g_qp.p, qstart.p, dp.p, qp.p);
    LAUNCH(k_pack, dim3(blocks(NA, 128)), dim3(128), 0, b, (const int*)active.p, qstart.p, dp.p, qp.p,
                       qin.p, (const double*)nullptr, (double*)nullptr);
    double econv = 1e-6 * o.accuracy, pconv = 2e-5 * o.accuracy;
    hactive = in_mask;
    hsplit.assign(NM, 0);
    std::vector<GemmJob> g1(NM), g2(NM), g3v(NM), gP(NM);
    std::vector<EigJob> ejobs(NM);
    std::vector<int> isdense(NM);
    bool anydense = false;
    mtau.ensure(NO);
    for (int m = 0; m < NM; m++) {
      const Mol& M = mols[m];
      long long o0 = M.mat0;
      int n = M.nao;
      isdense[m] = dense(n);
      anydense |= isdense[m] != 0;
      // the eigenproblem in the basis B: the Loewdin basis X at the start, then the previous
      // iteration's orbitals (A = B^T F B is then nearly diagonal and Jacobi needs few sweeps).
      // The dense solver does not profit from a nearly diagonal matrix: B = X throughout.
      double* B = isdense[m] ? X.p + o0 : Bm.p + o0;
      g1[m] = GemmJob{n, n, n, 0, 0, F.p + o0, B, T.p + o0, nullptr, n, n, n, 1.0, 0.0};   // T = F B
      g2[m] = GemmJob{n, n, n, 1, 0, B, T.p + o0, Am.p + o0, nullptr, n, n, n, 1.0, 0.0};  // A = B^T T
      // C = B U (mixed precision: U after the refinement steps, also in Z)
      g3v[m] = GemmJob{n, n, n, 0, 0, B, Z.p + o0, C.p + o0, nullptr, n, n, n, 1.0, 0.0};
      gP[m] = GemmJob{n, n, n, 0, 1, C.p + o0, C.p + o0, P.p + o0, focc.p + M.ao0, n, n, n, 1.0, 0.0};
      ejobs[m] = EigJob{n, Am.p + o0, Vw.p + o0, w.p + M.ao0, Z.p + o0};
    }
    std::vector<int> hkocc(NM);
    // dense molecules in mixed mode: the next eigenproblem by a full solve (all n vectors kept for
    // the refinement) or by refining the previous iteration's vectors
    // mixed and FP64-refine: the refinement of the previous orbitals (split for batched molecules,
    // dense refinement); mixed alone adds the FP32 Jacobi and the FP32 refinement update
    prec = o.precision;
    const bool mixed = o.precision != AXTB_PREC_FP64;
    const bool fp32 = o.precision == AXTB_PREC_MIXED;
    std::vector<int> dfull(NM, 1), drefined(NM, 0);
    dense_pfp.resize(NM, 0);
    for (int m = 0; m < NM; m++)
      if (in_mask[m]) dense_pfp[m] = 0;
    HIP_OK(hipMemcpyAsync(Bm.p, X.p, sizeof(double) * tot_mat, hipMemcpyDeviceToDevice, st));
    // The fixed job lists of this SCF (products, refinement steps, eigensolves) hold pointers into
    // the engine's buffers and per-molecule flags only: an SCF of the same batch, mask and
    // precision over the same buffers reuses the lists already on the device.
    rflags.ensure(4 * (size_t)NM);
    std::vector<long long> sig;
    bool cached = false;
    if (!t_arena) {
      sig = {batch_gen, NM, (long long)o.precision};
      for (const void* q : {(const void*)F.p, (const void*)T.p, (const void*)Am.p, (const void*)Bm.p, (const void*)C.p,
                            (const void*)P.p, (const void*)Z.p, (const void*)Vw.p, (const void*)Wm.p, (const void*)w.p,
                            (const void*)focc.p, (const void*)X.p, (const void*)active.p, (const void*)rflags.p})
        sig.push_back((long long)(intptr_t)q);
      sig.insert(sig.end(), in_mask.begin(), in_mask.end());
      cached = sig == list_sig;
    }
    list_sig.clear();  // set again once the lists below are on the device
    std::vector<GemmJob>(&sl)[4] = sl_;
    size_t(&sloff)[4] = sloff_;
    // the products of every iteration for the batched molecules of this SCF (T = F B, A = B^T T,
    // C = B U, P = C f C^T), uploaded once; each runs while its molecule is active
    if (!cached) {
      for (auto& l : sl) l.clear();
      for (int m : by_size) {
        if (!in_mask[m] || isdense[m]) continue;
        GemmJob j[4] = {g1[m], g2[m], g3v[m], gP[m]};
        j[1].lower = j[3].lower = sym_lower(mols[m].nao, false);
        for (int k = 0; k < 4; k++) {
          j[k].run = active.p + m;
          sl[k].push_back(j[k]);
        }
      }
      std::vector<GemmJob> all;
      for (int k = 0; k < 4; k++) {
        sloff[k] = all.size();
        all.insert(all.end(), sl[k].begin(), sl[k].end());
      }
      gmeta.clear();
      sjobs.upload(all, st, &stage);
    }
    auto gemm_static = [&](int k) {
      if (!sl[k].empty()) gemm(sl[k], sjobs.p + sloff[k]);
    };
    if (mixed && !cached) {  // the refinement's step lists of every batched molecule of this SCF, uploaded once
      std::vector<EigJob> ball;
      std::vector<int> ballm;
      for (int m : by_size)
        if (in_mask[m] && !isdense[m]) {
          ball.push_back(ejobs[m]);
          ballm.push_back(m);
        }
      refine_prepare(ball, ballm);
    }
    // An SCF without dense-solver molecules runs every launch of an iteration over fixed job lists
    // with device flags (the eigensolves too), so that the host enqueues an iteration while the
    // previous one runs and reads the active mask one iteration late: a batch that has converged
    // runs one more iteration of launches that all skip. With profiling on, the mask is read at the
    // end of each iteration (the per-launch work counts then see the exact set).
    bool anyd = false, anym = false;
    for (int m = 0; m < NM; m++) {
      anyd |= in_mask[m] && isdense[m];
      anym |= in_mask[m] != 0;
    }
    const bool devloop = !anyd && !t_arena;
    std::vector<EigJob> dall;
    std::vector<int> dallm;
    if (devloop) {
      for (int m : by_size)
        if (in_mask[m]) {
          dall.push_back(ejobs[m]);
          dallm.push_back(m);
        }
      const int ND = (int)dall.size();
      splitok.ensure(NM);
      dpfp.ensure(NM);
      LAUNCH(k_zero_u32, dim3(blocks(4LL * NM, 256)), dim3(256), 0, (unsigned*)rflags.p, 4 * NM);
      LAUNCH(k_zero_u32, dim3(blocks(NM, 256)), dim3(256), 0, (unsigned*)splitok.p, NM);
      if (!cached) {
        dlist.upload(dallm, st, &stage);
        std::vector<EigJob> j64 = dall;
        for (int q = 0; q < ND; q++) j64[q].run = (mixed ? rflags.p + 2 * (size_t)NM : active.p) + dallm[q];
        eigh_plan(ep64, j64, false, -1, 1.0f, &dallm);
        if (fp32) {
          std::vector<EigJob> j32 = dall;
          for (int q = 0; q < ND; q++) j32[q].run = rflags.p + NM + dallm[q];
          eigh_plan(ep32, j32, true, 60, 1.0f, nullptr);
        }
        if (mixed) {
          rg_jobs.upload(dall, st, &stage);
          rg_mol.upload(dallm, st, &stage);
        }
      }
      if (pin_cap < 2 * (size_t)NM) {
        if (pin_act) HIP_OK(hipHostFree(pin_act));
        pin_act = nullptr;
        HIP_OK(hipHostMalloc((void**)&pin_act, 2 * sizeof(int) * NM, hipHostMallocDefault));
        pin_cap = 2 * (size_t)NM;
      }
    }
    list_sig = sig;
    int it = 0;
    for (it = 1; it <= o.max_iter; it++) {
      std::vector<GemmJob> a1, a2;  // the dense molecules' products (the batched ones: gemm_static)
      std::vector<EigJob> ea;
      std::vector<int> eam;  // molecule of each entry of ea
      bool any_active = false;
      if (devloop) {
        // the mask after iteration it - 2 (its slot is it's parity); iteration it - 1 is queued
        if (!prof.on && it >= 3) {
          HIP_OK(hipEventSynchronize(ev_act[it & 1]));
          std::memcpy(hactive.data(), pin_act + (size_t)(it & 1) * NM, sizeof(int) * NM);
        }
        for (int m = 0; m < NM; m++) any_active |= hactive[m] != 0;
        if (it == 1) any_active = anym;
      } else {
        for (int m = 0; m < NM; m++)
          if (hactive[m]) {
            any_active = true;
            if (isdense[m]) {
              a1.push_back(g1[m]);
              a2.push_back(g2[m]);
            } else {
              ea.push_back(ejobs[m]);
              eam.push_back(m);
            }
          }
      }
      if (!any_active) break;
      if (it > 1 && mixer == MIX_DIIS) {
        LAUNCH(k_diis, dim3(NM), dim3(256), 0, b, (const int*)active.p, qin.p, (const double*)dq.p, dhq.p, dhr.p,
               dbq.p, dbr.p, dbas.p, (long long)tot_vec, ddst.p, dist.p, status.p);
      } else if (it > 1) {
        // history slots the kernel reads and writes in this iteration, and the order of its system
        const int ld = std::max(1, std::min(cap, it - 2));
        ensure_hist(ld);
        const size_t lds_small = (16 + 2 * (size_t)ld + (size_t)ld * ld) * sizeof(double);
        const bool in_lds = lds_small <= 64 * 1024;
        if (!in_lds) gbeta.ensure((size_t)NM * ld * ld);
        const size_t bshm = in_lds ? lds_small : (16 + 2 * (size_t)ld) * sizeof(double);
        if (bshm > 48 * 1024)
          HIP_OK(hipFuncSetAttribute((const void*)k_broyden, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bshm));
        LAUNCH(k_broyden, dim3(NM), dim3(256), bshm, b, active.p, cap, ld, o.damping, qin.p, dq.p, qlast.p, dqlast.p,
               hdf.p, hu.p, (long long)tot_vec, hom.p, ham.p, NM, in_lds ? (double*)nullptr : gbeta.p, itc.p,
               status.p);
      }
      LAUNCH(k_unpack, dim3(blocks(NA, 128)), dim3(128), 0, b, active.p, qin.p, qsh.p, qat.p, dp.p,
                         qp.p);
      // potentials of the input moments (xtb.rs:316-340)
      LAUNCH(k_shell_pot, dim3(blocks(NS, 4)), dim3(256), 0, b, active.p, gam.p, qsh.p, vsh.p, 1);
      LAUNCH(k_d4_weights, dim3(blocks(NA, 128)), dim3(128), 0, b, active.p, cnd4.p, qat.p, gw.p,
                         dgw.p, 0);
      if (!o.dispersion) fill(gw.p, MAXREF * (long long)NA);
      LAUNCH(k_atom_pot, dim3(blocks(NA, 4)), dim3(256), 0, b, active.p, mrad.p, qat.p, dp.p, qp.p,
                         gw.p, dgw.p, vat.p, vdp.p, vqp.p);
      if (solv_model != AXTB_SOLV_NONE)  // the solvation potential B q of the input charges
        LAUNCH(k_solv_pot, dim3(blocks(NA, 4)), dim3(256), 0, b, active.p, solv_dev(), (const double*)qat.p, vat.p);
      // HBM bytes of the streaming kernels: Fock reads the lower triangles of S, H0, D, Q and writes
      // F (11 n (n + 1) / 2 + n^2 doubles); Mulliken reads P, S, D, Q, H0 (12 n^2 doubles)
      double sbytes = 0.0, fbytes = 0.0;
      for (int m = 0; m < NM; m++)
        if (hactive[m]) {
          const double n = mols[m].nao;
          sbytes += 12.0 * 8.0 * n * n;
          fbytes += 8.0 * (11.0 * n * (n + 1.0) / 2.0 + n * n);
        }
      LAUNCHW(fbytes, k_fock, dim3(n_mat_tiles), dim3(256), 0, b, active.p, (const int2*)mat_tiles.p,
                         pool.p, stride, vsh.p, vat.p, vdp.p, vqp.p, F.p);
      // generalized eigenproblem in the Loewdin basis
      gemm_static(0);
      gemm(a1);
      gemm_static(1);  // A = B^T (F B) is symmetric: lower blocks
      for (auto& j : a2) j.lower = sym_lower(j.n, dense(j.n));
      gemm(a2);
      mirror_dense(a2);
      if (mixed) {
        // mixed: the FP32 Jacobi, then FP64 Ogita-Aishima steps U' = U (I + E) from S = U^T A U and
        // G = U^T U (U in Z, A in Am; S -> F, G -> Vw, I + E -> Wm, U' -> T or Z).
        // Molecules whose previous occupations were all 0 or 2 and whose last residual is below the
        // gate (decided on the device at the end of the previous iteration and read with the active
        // mask): the previous orbitals (A nearly block diagonal) refined for the occupied / empty
        // split alone, from U = I, no Jacobi; P needs no more, W is then P F P / 2 and the orbital
        // energies come from the blocks at the end. The others: the FP32 Jacobi and the full
        // refinement (FP64-refine: the FP64 Jacobi). A refinement that does not converge: the FP64
        // Jacobi. No host read in between: per-molecule flags (run, fail, in_t, split) set on the
        // device and launches over fixed job lists.
        const bool use_split = it >= split_from && split_refine;
        if (devloop) {
          const int ND = (int)dall.size();
          const int* f_split = rflags.p;
          const int* f_full = rflags.p + NM;
          int* f_fail = rflags.p + 2 * (size_t)NM;
          int* f_int = rflags.p + 3 * (size_t)NM;
          LAUNCH(k_iter_flags, dim3(blocks(ND, 256)), dim3(256), 0, ND, (const int*)dlist.p, (const int*)active.p,
                 (const int*)splitok.p, use_split ? 1 : 0, fp32 ? 1 : 0, NM, rflags.p);
          // the FP32 Jacobi of the molecules on the full path on st4 beside the split refinement of
          // the others (their buffers are disjoint), joined before the full refinement
          const bool beside = fp32 && use_split && jacobi_streams;
          if (beside) {
            HIP_OK(hipEventRecord(ev_fork4, st));
            HIP_OK(hipStreamWaitEvent(st4, ev_fork4, 0));
            eigh_launch(ep32, st4);
            HIP_OK(hipEventRecord(ev_join4, st4));
          }
          if (use_split) refine_run(f_split, f_int, f_fail, ocls.p);
          if (fp32) {
            if (beside)
              HIP_OK(hipStreamWaitEvent(st, ev_join4, 0));
            else
              eigh_launch(ep32);
            refine_run(f_full, f_int, f_fail, nullptr);
          }
          eigh_launch(ep64);
          LAUNCH(k_refine_gather, dim3((unsigned)ND), dim3(256), 0, (const EigJob*)rg_jobs.p, (const int*)rg_mol.p,
                 (long long)(T.p - Z.p), (const int*)f_int, (const int*)f_fail, f_split, dpfp.p, (const int*)active.p);
        }
        std::vector<EigJob> ef;
        int nsplit = 0;
        std::vector<int> hflags(4 * (size_t)NM, 0);  // split | full (FP32 path) | fail | in_t
        for (size_t q = 0; q < ea.size(); q++) {
          const int m = eam[q];
          const bool sp = use_split && hsplit[m];
          if (sp)
            nsplit++;
          else
            ef.push_back(ea[q]);
          hflags[m] = sp ? 1 : 0;
          hflags[(size_t)NM + m] = (!sp && fp32) ? 1 : 0;
          hflags[2 * (size_t)NM + m] = (!sp && !fp32) ? 1 : 0;
        }
        if (!devloop && !ea.empty()) {
          rflags.upload(hflags, st, &stage);
          const int* f_split = rflags.p;
          const int* f_full = rflags.p + NM;
          int* f_fail = rflags.p + 2 * (size_t)NM;
          int* f_int = rflags.p + 3 * (size_t)NM;
          if (nsplit) refine_run(f_split, f_int, f_fail, ocls.p);
          if (fp32 && !ef.empty()) {
            eigh(ef, true, 60);
            refine_run(f_full, f_int, f_fail, nullptr);
          }
          // not converged by a refinement, and FP64-refine's full solves: the FP64 Jacobi (vectors in Z)
          std::vector<EigJob> redo = ea;
          for (size_t q = 0; q < redo.size(); q++) redo[q].run = f_fail + eam[q];
          eigh(redo, false, -1, 1.0f, &eam);
          // C = B U reads Z: the refined vectors that ended in T are copied there
          dpfp.ensure(NM);
          rg_jobs.upload(ea, st, &stage);
          rg_mol.upload(eam, st, &stage);
          LAUNCH(k_refine_gather, dim3((unsigned)ea.size()), dim3(256), 0, (const EigJob*)rg_jobs.p,
                 (const int*)rg_mol.p, (long long)(T.p - Z.p), (const int*)f_int, (const int*)f_fail, f_split,
                 dpfp.p, (const int*)nullptr);
        }
      } else if (devloop) {
        eigh_launch(ep64);
      } else {
        eigh(ea, false, -1, 1.0f, &eam);
      }
      for (int m = 0; m < NM; m++) {
        if (!hactive[m] || !isdense[m]) continue;
        drefined[m] = mixed && dense_refine && !dfull[m];
        if (drefined[m]) {
          // the previous iteration's vectors refined for the new matrix (classes from its
          // occupations); not converged in dense_steps steps: a full solve instead
          const int n = mols[m].nao;
          dcls.ensure(n);
          dnfrac.ensure(1);
          LAUNCH(k_zero_u32, dim3(1), dim3(64), 0, (unsigned*)dnfrac.p, 1);
          LAUNCH(k_occ_class, dim3(blocks(n, 256)), dim3(256), 0, (const double*)focc.p + mols[m].ao0, n, dcls.p,
                 dnfrac.p);
          int nfrac = 0;
          HIP_OK(hipMemcpyAsync(&nfrac, dnfrac.p, sizeof(int), hipMemcpyDeviceToHost, st));
          HIP_OK(hipStreamSynchronize(st));
          const double emax = nfrac ? 1.0 : refine_big(ejobs[m], dcls.p);
          if (!(emax <= dense_accept)) {
            drefined[m] = 0;
            prof.count("refine_dense_fallbacks", 1, 1.0);
          }
        }
        if (!drefined[m]) dense_values(ejobs[m].A, ejobs[m].n, ejobs[m].w, mtau.p + mols[m].ao0, ejobs[m].V);
        dfull[m] = it < dense_full ? 1 : 0;
        dense_pfp[m] = drefined[m];
      }
      gemm_static(2);  // C = B U
      HIP_OK(hipMemcpyAsync(Bm.p, C.p, sizeof(double) * tot_mat, hipMemcpyDeviceToDevice, st));
      LAUNCH(k_fermi, dim3(NM), dim3(256), 0, b, active.p, w.p, focc.p, fwork.p, fst.p, kocc.p,
                         status.p);
      // dense molecules: only the kocc orbitals with occupation > 1e-16 are formed, C = X U
      if (anydense) {
        HIP_OK(hipMemcpyAsync(hkocc.data(), kocc.p, sizeof(int) * NM, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
      }
      std::vector<GemmJob> a4, a5;
      for (int m = 0; m < NM; m++) {
        if (!hactive[m] || !isdense[m]) continue;
        GemmJob p = gP[m];
        {
          const int n = mols[m].nao, k = hkocc[m];
          GemmJob c = g3v[m];
          if (!drefined[m]) {
            // mixed: all vectors, the start of the next iteration's refinement (refine_big reads and
            // leaves the vectors in Z)
            const bool keep_all = mixed && dense_refine && it >= dense_full;  // the next iteration refines
            dense_vectors(ejobs[m].A, n, mtau.p + mols[m].ao0, ejobs[m].V, keep_all ? n : k, ejobs[m].Z);
          }
          c.B = ejobs[m].Z;
          c.n = k;
          a5.push_back(c);
          p.k = k;
        }
        a4.push_back(p);
      }
      gemm(a5);
      gemm_static(3);  // P = C diag(f) C^T
      for (auto& j : a4) j.lower = sym_lower(j.n, dense(j.n));
      gemm(a4);
      mirror_dense(a4);
      LAUNCHW(sbytes, k_mulliken_cols, dim3(blocks(NO, 4)), dim3(256), 0, b, active.p, pool.p, stride, P.p,
                         cols.p);
      LAUNCH(k_mulliken_atoms, dim3(blocks(NA, 128)), dim3(128), 0, b, active.p, cols.p, qsh_o.p,
                         qat_o.p, dp_o.p, qp_o.p, eband.p);
      LAUNCH(k_pack, dim3(blocks(NA, 128)), dim3(128), 0, b, active.p, qsh_o.p, dp_o.p, qp_o.p,
                         qout.p, (const double*)qin.p, dq.p);
      // energy of the output moments
      LAUNCH(k_shell_pot, dim3(blocks(NS, 4)), dim3(256), 0, b, active.p, gam.p, qsh_o.p, gq.p, 0);
      LAUNCH(k_d4_weights, dim3(blocks(NA, 128)), dim3(128), 0, b, active.p, cnd4.p, qat_o.p, gw.p,
                         dgw.p, 0);
      if (!o.dispersion) fill(gw.p, MAXREF * (long long)NA);
      LAUNCH(k_energy_atoms, dim3(blocks(NA, 4)), dim3(256), 0, b, active.p, mrad.p, gq.p, qsh_o.p,
                         qat_o.p, dp_o.p, qp_o.p, gw.p, dgw.p, eband.p, fst.p, eel.p, eterm.p);
      if (solv_model != AXTB_SOLV_NONE)  // q (B q) / 2 of the output charges
        LAUNCH(k_solv_energy, dim3(blocks(NA, 4)), dim3(256), 0, b, active.p, solv_dev(), (const double*)qat_o.p,
               eel.p, eterm.p);
      LAUNCH(k_converge, dim3(NM), dim3(256), 0, b, active.p, eel.p, dq.p, fst.p, econv, pconv, mst.p,
                         status.p);
      // mixed: which batched molecules may refine for the occupied / empty split in the next
      // iteration (occupation classes of this iteration's orbitals, this iteration's residual),
      // read with the active mask: the one host read of an iteration
      const bool next_split = mixed && split_refine && it + 1 >= split_from;
      if (next_split && devloop) {
        const int ND = (int)dall.size();
        ocls.ensure(NO);
        onfrac.ensure(NM);
        LAUNCH(k_occ_class_batch, dim3((unsigned)ND), dim3(256), 0, b, (const int*)dlist.p, (const double*)focc.p,
               ocls.p, onfrac.p);
        LAUNCH(k_split_ok, dim3(blocks((long long)ND, 256)), dim3(256), 0, (const int*)dlist.p, ND,
               (const int*)onfrac.p, (const double*)mst.p, split_residual, splitok.p);
      }
      if (devloop) {
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(pin_act + (size_t)(it & 1) * NM, active.p, sizeof(int) * NM, hipMemcpyDeviceToHost, st));
        HIP_OK(hipEventRecord(ev_act[it & 1], st));
        if (prof.on) {
          HIP_OK(hipEventSynchronize(ev_act[it & 1]));
          std::memcpy(hactive.data(), pin_act + (size_t)(it & 1) * NM, sizeof(int) * NM);
          bool any = false;
          for (int m = 0; m < NM; m++) any |= hactive[m] != 0;
          if (!any) break;
        }
        continue;
      }
      if (next_split) {
        std::vector<int> ml;
        for (int m = 0; m < NM; m++)
          if (hactive[m] && !isdense[m]) ml.push_back(m);
        if (!ml.empty()) {
          dmlist.upload(ml, st, &stage);
          ocls.ensure(NO);
          onfrac.ensure(NM);
          splitok.ensure(NM);
          LAUNCH(k_occ_class_batch, dim3(ml.size()), dim3(256), 0, b, (const int*)dmlist.p, (const double*)focc.p,
                 ocls.p, onfrac.p);
          LAUNCH(k_split_ok, dim3(blocks((long long)ml.size(), 256)), dim3(256), 0, (const int*)dmlist.p,
                 (int)ml.size(), (const int*)onfrac.p, (const double*)mst.p, split_residual, splitok.p);
          HIP_OK(hipMemcpyAsync(hsplit.data(), splitok.p, sizeof(int) * NM, hipMemcpyDeviceToHost, st));
        }
      }
      HIP_OK(hipGetLastError());
      HIP_OK(hipMemcpyAsync(hactive.data(), active.p, sizeof(int) * NM, hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
      bool any = false;
      for (int m = 0; m < NM; m++) any |= hactive[m] != 0;
      if (!any) break;
    }
    iters_ = it;
    if (mixed && dpfp.p) {  // batched molecules: whether their last iteration was a converged split refinement
      std::vector<int> hp(NM);
      HIP_OK(hipMemcpyAsync(hp.data(), dpfp.p, sizeof(int) * NM, hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
      for (int m = 0; m < NM; m++)
        if (in_mask[m] && !isdense[m]) dense_pfp[m] = hp[m];
    }
    // refined dense molecules: the orbital energies of the last iteration's matrix (the refinement
    // leaves Rayleigh quotients inside the occupied and empty blocks; P, W and the energy do not
    // use them) as the eigenvalues of the occupied and the empty block of S = U^T A U from the last
    // refinement step (in F; the first kocc columns are the occupied ones, and the coupling between
    // the blocks is below the acceptance bound, so the blocks' eigenvalues are A's to second order
    // in it), each by an eigenvalue-only FP64 dense solve
    bool anypfp = false;
    for (int m = 0; m < NM; m++) anypfp |= dense_pfp[m] != 0 && in_mask[m];
    if (!mixed || !anypfp) return;
    std::vector<int> hk(NM);
    HIP_OK(hipMemcpyAsync(hk.data(), kocc.p, sizeof(int) * NM, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    // batched molecules refined for the occupied / empty split in their last iteration (always, as
    // their orbital energies are Rayleigh quotients otherwise): both blocks of S of every one of
    // them, by the batched FP64 Jacobi (blocks in T, vectors to Vw, work in C: free now)
    {
      std::vector<EigJob> bj;
      std::vector<BlockCopy> bc;
      std::vector<int> bmol;
      for (int m = 0; m < NM; m++) {
        if (!dense_pfp[m] || !in_mask[m] || dense(mols[m].nao)) continue;
        const int n = mols[m].nao, k = hk[m];
        const long long o0 = mols[m].mat0, a0 = mols[m].ao0;
        long long off = 0;
        for (int blk = 0; blk < 2; blk++) {
          const int s0 = blk == 0 ? 0 : k, nb = blk == 0 ? k : n - k;
          if (nb <= 0) continue;
          bc.push_back(BlockCopy{F.p + o0 + s0 + (long long)s0 * n, T.p + o0 + off, nb, n});
          bj.push_back(EigJob{nb, T.p + o0 + off, nullptr, w.p + a0 + s0, nullptr});  // eigenvalues only
          bmol.push_back(m);
          off += (long long)nb * nb;
        }
      }
      if (!bc.empty()) {
        dbcopy.upload(bc, st, &stage);
        LAUNCH(k_copy_blocks, dim3(bc.size()), dim3(256), 0, (const BlockCopy*)dbcopy.p);
      }
      eigh(bj, false, -1, 1.0f, &bmol);
    }
    if (exact_levels) {
      for (int m = 0; m < NM; m++) {
        if (!dense_pfp[m] || !in_mask[m] || !dense(mols[m].nao)) continue;
        const int n = mols[m].nao, k = hk[m];
        const long long o0 = mols[m].mat0, a0 = mols[m].ao0;
        for (int blk = 0; blk < 2; blk++) {
          const int s0 = blk == 0 ? 0 : k, nb = blk == 0 ? k : n - k;
          if (nb <= 0) continue;
          HIP_OK(hipMemcpy2DAsync(T.p + o0, sizeof(double) * nb, F.p + o0 + s0 + (long long)s0 * n, sizeof(double) * n,
                                  sizeof(double) * nb, nb, hipMemcpyDeviceToDevice, st));
          dense_values(T.p + o0, nb, w.p + a0 + s0, mtau.p + a0 + s0, Vw.p + o0);
        }
      }
    }
    // the Fermi level and -TS of the refined molecules from their exact levels (the occupations
    // stay 0 or 2), into the iteration record
    std::vector<int> hp(NM, 0);
    for (int m = 0; m < NM; m++) hp[m] = in_mask[m] && dense_pfp[m] && (exact_levels || !dense(mols[m].nao)) ? 1 : 0;
    dmlist.upload(hp, st, &stage);
    LAUNCH(k_fermi, dim3(NM), dim3(256), 0, b, (const int*)dmlist.p, w.p, focc.p, fwork.p, fst.p, kocc.p, status.p);
    HIP_OK(hipMemcpy2DAsync(mst.p + 3, 8 * sizeof(double), fst.p, 2 * sizeof(double), 2 * sizeof(double), NM,
                            hipMemcpyDeviceToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
  }
  bool exact_levels = true;  // mixed, dense: exact orbital energies at the end (AXTB_HIP_EXACT_LEVELS)
  int prec = AXTB_PREC_FP64;  // precision of the running SCF

  // ------------------------------------------------------------------------------ gradient
  std::vector<int> hok;
  DBuf<int> gokd, gpfd;  // the gradient's molecule masks (converged without errors; W = P F P / 2)

  void gradient() {
    Batch b = batch();
    const int NA = nat(), NM = nmol();
    const long long stride = tot_mat;
    // molecules that converged without errors
    std::vector<double> hm(8 * (size_t)NM);
    std::vector<int> hst(NM);
    HIP_OK(hipMemcpyAsync(hm.data(), mst.p, sizeof(double) * 8 * NM, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(hst.data(), status.p, sizeof(int) * NM, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    hok.assign(NM, 0);
    std::vector<GemmJob> gW, gF, gW2;
    std::vector<int> hpfp(NM, 0);
    bool anypfp = false;
    int naomax = 0;
    std::vector<int> hk(NM);
    HIP_OK(hipMemcpyAsync(hk.data(), kocc.p, sizeof(int) * NM, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    for (int m : by_size) {  // the product lists largest first
      hok[m] = (hm[8 * m + 6] == 1.0 && hst[m] == 0) ? 1 : 0;
      if (!hok[m]) continue;
      const Mol& M = mols[m];
      int n = M.nao;
      if (m < (int)dense_pfp.size() && dense_pfp[m]) {
        // refined occupied subspace (occupations 0 or 2): W = P F P / 2 with F rebuilt from the last
        // input potentials (the refinement used the F pool as work space)
        hpfp[m] = 1;
        anypfp = true;
        gF.push_back(GemmJob{n, n, n, 0, 0, F.p + M.mat0, P.p + M.mat0, T.p + M.mat0, nullptr, n, n, n, 1.0, 0.0});
        gW2.push_back(GemmJob{n, n, n, 0, 0, P.p + M.mat0, T.p + M.mat0, Wm.p + M.mat0, nullptr, n, n, n, 0.5, 0.0});
        naomax = std::max(naomax, n);
        continue;
      }
      // dense molecules hold only the first kocc orbitals (the others have occupation <= 1e-16)
      int k = dense(n) ? hk[m] : n;
      gW.push_back(GemmJob{n, n, k, 0, 1, C.p + M.mat0, C.p + M.mat0, Wm.p + M.mat0, fwork.p + M.ao0, n, n, n, 1.0, 0.0});
    }
    if (anypfp) {
      DBuf<int>& pfd = gpfd;
      pfd.upload(hpfp, st, &stage);
      LAUNCH(k_fock, dim3(n_mat_tiles), dim3(256), 0, b, (const int*)pfd.p, (const int2*)mat_tiles.p, pool.p,
             stride, vsh.p, vat.p, vdp.p, vqp.p, F.p);
      gemm(gF);
      for (auto& j : gW2) j.lower = 1;  // W = P (F P) / 2 is symmetric
      gemm(gW2);
      mirror_dense(gW2);
      HIP_OK(hipStreamSynchronize(st));
    }
    DBuf<int>& okd = gokd;
    okd.upload(hok, st, &stage);
    // focc * eps into fwork (k_fermi's work is no longer needed)
    LAUNCH(k_pairwise, dim3(std::min<long long>(blocks(naos(), 256), 4096)), dim3(256), 0, (const double*)focc.p,
           (const double*)w.p, (long long)naos(), 1, fwork.p);
    for (auto& j : gW) j.lower = sym_lower(j.n, dense(j.n));
    gemm(gW);
    mirror_dense(gW);
    // gradient = repulsion + three-body (setup) + pair terms + CN chains
    LAUNCH(k_pairwise, dim3(std::min<long long>(blocks(3LL * NA, 256), 4096)), dim3(256), 0, (const double*)grep.p,
           (const double*)g3.p, 3LL * NA, 0, grad.p);
    LAUNCH(k_d4_weights, dim3(blocks(NA, 128)), dim3(128), 0, b, (const int*)okd.p, cnd4.p, qat_o.p,
                       gw.p, dgw.p, 1);
    int natmax = 0;
    for (auto& M : mols) natmax = std::max(natmax, M.nat);
    // H0 / overlap terms per atom pair by shell-pair type (hg in atmP: the ATM table is used up);
    // the lists of group g (molecules m % H0_GROUPS == g) hold up to the group's sum of nat^2
    {
      std::vector<long long> go(H0_GROUPS, 0);
      for (int m = 0; m < NM; m++)
        if (m % H0_GROUPS + 1 < H0_GROUPS) go[m % H0_GROUPS + 1] += (long long)mols[m].nat * mols[m].nat;
      for (int g = 1; g < H0_GROUPS; g++) go[g] += go[g - 1];
      h0goff.upload(go, st, &stage);
    }
    h0cnt.ensure(NKEYS * H0_GROUPS);
    h0list.ensure(NKEYS * (size_t)tot_pair);
    LAUNCH(k_zero_u32, dim3(blocks(NKEYS * H0_GROUPS, 256)), dim3(256), 0, (unsigned*)h0cnt.p, NKEYS * H0_GROUPS);
    LAUNCH(k_h0_keys, dim3(blocks((long long)natmax * natmax, 64), NM), dim3(64), 0, b, (const int*)okd.p, h0cnt.p,
           h0list.p, (long long)tot_pair, (const long long*)h0goff.p);
    fill(atmP.p, 5 * (long long)tot_pair);
    std::vector<int> hcnt(NKEYS * H0_GROUPS);
    HIP_OK(hipMemcpyAsync(hcnt.data(), h0cnt.p, sizeof(int) * hcnt.size(), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    auto h0key = [&](int k, auto kern) {
      int mx = 0;
      for (int g = 0; g < H0_GROUPS; g++) mx = std::max(mx, hcnt[k * H0_GROUPS + g]);
      if (mx <= 0) return;
      ProfScope ps_(&prof, st, "k_h0_grad", 0.0);
      hipLaunchKernelGGL(kern, dim3(blocks(mx, 64), H0_GROUPS), dim3(64), 0, st, b,
                         (const int2*)h0list.p + (long long)k * tot_pair, (const int*)h0cnt.p + k * H0_GROUPS,
                         (const long long*)h0goff.p, P.p, Wm.p, se.p, dsedcn.p, vsh.p, vat.p, vdp.p, vqp.p, atmP.p,
                         (long long)tot_pair);
    };
    h0key(0, k_h0_grad_key<0, 0>);
    h0key(1, k_h0_grad_key<0, 1>);
    h0key(2, k_h0_grad_key<0, 2>);
    h0key(3, k_h0_grad_key<1, 0>);
    h0key(4, k_h0_grad_key<1, 1>);
    h0key(5, k_h0_grad_key<1, 2>);
    h0key(6, k_h0_grad_key<2, 0>);
    h0key(7, k_h0_grad_key<2, 1>);
    h0key(8, k_h0_grad_key<2, 2>);
    LAUNCH(k_pair_grad, dim3(blocks(NA, 4)), dim3(256), 0, b, (const int*)okd.p, pool.p, stride, P.p, Wm.p, se.p,
           dsedcn.p, vsh.p, vat.p, vdp.p, vqp.p, qsh_o.p, qat_o.p, dp_o.p, qp_o.p, mrad.p, dmrdcn.p, gw.p, dgw.p,
           dcn3.p, (const double*)atmP.p, (long long)tot_pair, grad.p, dedcn.p, dedcn4.p);
    LAUNCH(k_cn_chain, dim3(blocks(NA, 4)), dim3(256), 0, b, (const int*)okd.p, dedcn.p, dedcn4.p,
                       grad.p);
    if (solv_model != AXTB_SOLV_NONE) {
      // solvation at the output charges: kernel and Born-radius chain, surface (with the H-bond
      // term), ALPB shape (lib.rs:688-709, 627-636)
      const SolvDev S = solv_dev();
      LAUNCH(k_solv_grad_pairs, dim3(blocks(NA, 4)), dim3(256), 0, b, (const int*)okd.p, S, (const double*)qat_o.p,
             grad.p);
      LAUNCH(k_solv_grad_psi, dim3(blocks(NA, 4)), dim3(256), 0, b, (const int*)okd.p, S, grad.p);
      LAUNCH(k_solv_grad_sasa, dim3(NA), dim3(64), 0, b, (const int*)okd.p, S, (const double*)qat_o.p, grad.p);
      if (S.alpb) LAUNCH(k_solv_grad_adet, dim3(NM), dim3(256), 0, b, (const int*)okd.p, S, (const double*)qat_o.p, grad.p);
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(st));
  }

  // ------------------------------------------------------------------------------ results
  void collect(const axtb_opts& o, std::vector<MolOut>& out) {
    const int NA = nat(), NS = nshell(), NO = naos(), NM = nmol();
    auto get = [&](DBuf<double>& d, size_t n) {
      std::vector<double> v(n);
      if (n) HIP_OK(hipMemcpyAsync(v.data(), d.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
      return v;
    };
    auto m8 = get(mst, 8 * (size_t)NM);
    auto er = get(erep, NA), e3h = get(e3, NA), eelh = get(eel, NA), et = get(eterm, 8 * (size_t)NA);
    auto qa = get(qat_o, NA), qs = get(qsh_o, NS), d = get(dp_o, 3 * (size_t)NA), q = get(qp_o, 6 * (size_t)NA);
    auto e = get(w, NO);
    auto ef = solv_model != AXTB_SOLV_NONE ? get(sv_efix, NA) : std::vector<double>(NA, 0.0);
    auto g = grad_done ? get(grad, 3 * (size_t)NA) : std::vector<double>();
    std::vector<int> hst(NM), hef(NM);
    HIP_OK(hipMemcpyAsync(hst.data(), status.p, sizeof(int) * NM, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (out.size() < mols.size() && !mols.empty()) throw Fail{AXTB_ERR_INPUT, "result list shorter than the batch"};
    for (int m = 0; m < NM; m++) {
      const Mol& M = mols[m];
      if (map_use[m] >= (int)out.size()) throw Fail{AXTB_ERR_INPUT, "result list shorter than the batch"};
      MolOut& r = out[map_use[m]];
      double* s = &m8[8 * m];
      r.iterations = (int)s[5];
      // the SCF report (MixerPlan::used, scf/mixing.rs:150-157): on failure the iterations of all
      // attempts (xtb.rs:640-645)
      const Attempts& A = m < (int)att.size() ? att[m] : Attempts();
      r.nattempts = A.n;
      for (int k = 0; k < 2; k++) {
        r.att_iter[k] = A.iter[k];
        r.att_conv[k] = A.conv[k];
      }
      r.mixer_used = (o.mixer == AXTB_MIXER_AUTO && A.n > 1) ? 2 : (A.n ? A.kind[A.n - 1] : 0);
      if (s[6] != 1.0) r.iterations = A.iter[0] + (A.n > 1 ? A.iter[1] : 0);
      r.residual = s[2];
      r.energy_change = s[1];
      r.efermi = s[3];
      if (hst[m] & 32) {
        r.status = AXTB_ERR_NUMERICAL;
        r.msg = "solvation: more atoms than the surface kernel takes within the surface cutoff of one atom";
      } else if (hst[m] & 2) {
        r.status = AXTB_ERR_NUMERICAL;
        r.msg = "Fermi level not found";
      } else if (hst[m] & 4) {
        r.status = AXTB_ERR_NUMERICAL;
        r.msg = "non-finite energy";
      } else if (hst[m] & 16) {
        r.status = AXTB_ERR_NUMERICAL;
        r.msg = "the eigensolver did not converge";
      } else if (hst[m] & 8) {
        r.status = AXTB_ERR_NUMERICAL;
        r.msg = (A.n && A.kind[A.n - 1] == MIX_DIIS) ? "DIIS mixing failed" : "Broyden mixing failed";
      } else if (s[6] != 1.0) {
        r.status = AXTB_ERR_SCF_NOT_CONVERGED;
        r.msg = "SCF not converged in " + std::to_string(r.iterations) + " iterations";
      }
      double etot = 0.0, erp = 0.0, ed3 = 0.0, eelt = 0.0, esf = 0.0, t[6] = {0, 0, 0, 0, 0, 0};
      const int nt = solv_model != AXTB_SOLV_NONE ? 6 : 5;
      for (int a = M.at0; a < M.at0 + M.nat; a++) {
        etot += er[a] + e3h[a] + ef[a] + eelh[a];
        erp += er[a];
        ed3 += e3h[a];
        esf += ef[a];
        eelt += eelh[a];
        for (int k = 0; k < nt; k++) t[k] += et[8 * a + k];
      }
      r.energy = etot;
      // terms: repulsion dispersion electronic isotropic_es third_order anisotropic entropy band
      r.terms[0] = erp;
      r.terms[1] = ed3 + t[3];
      r.terms[2] = eelt;
      r.terms[3] = t[0];
      r.terms[4] = t[1];
      r.terms[5] = t[2];
      r.terms[6] = s[4];
      r.terms[7] = t[4];
      // the solvation free energy: its charge-independent part (also in the total beside the
      // repulsion) and its charge-dependent part (also in the electronic energy), xtb.rs:1094-1119
      r.terms[8] = esf + t[5];
      r.qat.assign(qa.begin() + M.at0, qa.begin() + M.at0 + M.nat);
      r.qsh.assign(qs.begin() + M.sh0, qs.begin() + M.sh0 + M.nsh);
      r.dp.assign(d.begin() + 3 * M.at0, d.begin() + 3 * (M.at0 + M.nat));
      r.qp.assign(q.begin() + 6 * M.at0, q.begin() + 6 * (M.at0 + M.nat));
      r.emo.assign(e.begin() + M.ao0, e.begin() + M.ao0 + M.nao);
      std::sort(r.emo.begin(), r.emo.end());
      // molecular dipole about the origin (xtb.rs:746-751), gap above the alpha HOMO (properties.rs:204-221)
      for (int k = 0; k < 3; k++) {
        r.dipole[k] = 0.0;
        for (int a = 0; a < M.nat; a++) r.dipole[k] += qa[M.at0 + a] * hxyz[3 * (M.at0 + a) + k] + d[3 * (M.at0 + a) + k];
      }
      {
        const double nel = std::max(M.nalpha, M.nbeta), fl = std::floor(nel);
        const double kk = nel - fl > 0.5 ? fl + 1.0 : fl;
        r.gap = (kk >= 1.0 && (int)kk < M.nao) ? r.emo[(int)kk] - r.emo[(int)kk - 1] : std::nan("");
      }
      if (grad_done && r.status == 0) r.grad.assign(g.begin() + 3 * M.at0, g.begin() + 3 * (M.at0 + M.nat));
    }
  }
};

}  // namespace axtb

// ================================================================================ C ABI
using axtb::Engine;

// Device memory of the no-allocation entry point (axtbhip_gfn2_engrad_noalloc): one array per GPU,
// held until the process ends. Host code never uses the name itself (on the host it is a stand-in,
// not a device address): each device's address comes from hipGetSymbolAddress at context creation.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
constexpr int NOALLOC_MAX_NAO = 256;  // the largest basis the no-allocation entry point takes
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {nullptr};  // per device: the address of `scratch`

// the device address of `scratch` on `device` (that device current), looked up once
void* scratch_address(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= 64) throw axtb::DeviceError("device index out of range");
  if (!g_scratch_ptr[device]) {
    HIP_OK(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !ptr)
      throw axtb::DeviceError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}
}  // namespace

struct axtbhip_context {
  std::unique_ptr<Engine> eng;
  std::string err;
  // the no-allocation path: its own engine (streams made here), its buffers carved from scratch
  std::unique_ptr<Engine> eng_na;
  axtb::Arena arena;
};

// A batch of single points; a molecule that started from given moments and did not converge runs
// once more from the fresh start (RetryLadder::warm_then_fresh, crates/base/src/method/retry.rs:150-158).
static void run_warm_then_fresh(Engine& eng, const std::vector<axtb::MolIn>& in, const axtb_opts& o,
                                std::vector<axtb::MolOut>& res) {
  eng.run(in, o, res);
  std::vector<int> again;
  for (size_t i = 0; i < in.size(); i++)
    if (in[i].gq && res[i].status == AXTB_ERR_SCF_NOT_CONVERGED) again.push_back((int)i);
  if (again.empty()) return;
  std::vector<axtb::MolIn> in2;
  for (int i : again) {
    in2.push_back(in[i]);
    in2.back().gq = in2.back().gdp = in2.back().gqp = nullptr;
  }
  std::vector<axtb::MolOut> res2;
  eng.run(in2, o, res2);
  for (size_t k = 0; k < again.size(); k++) res[again[k]] = res2[k];
}

static void fill_result(const axtb::MolOut& r, const axtb_mol& m, axtb_result* out) {
  out->status = r.status;
  out->energy = r.energy;
  out->iterations = r.iterations;
  out->nsh = r.nsh;
  out->nao = r.nao;
  out->efermi = r.efermi;
  out->residual = r.residual;
  out->energy_change = r.energy_change;
  for (int k = 0; k < 9; k++) out->terms[k] = r.terms[k];
  auto put = [](double* dst, const std::vector<double>& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(double));
  };
  put(out->gradient, r.grad);
  put(out->charges, r.qat);
  put(out->shell_charges, r.qsh);
  put(out->dipoles, r.dp);
  put(out->quadrupoles, r.qp);
  put(out->orbital_energies, r.emo);
  for (int k = 0; k < 3; k++) out->dipole[k] = r.dipole[k];
  out->gap = r.gap;
  out->mixer_used = r.mixer_used;
  out->nattempts = r.nattempts;
  for (int k = 0; k < 2; k++) {
    out->attempt_iterations[k] = r.att_iter[k];
    out->attempt_converged[k] = r.att_conv[k];
  }
  out->element = r.element;
  (void)m;
}

extern "C" {

const char* axtbhip_version(void) { return "axtb-hip 0.1.0 (GFN2-xTB, gfx950)"; }

axtbhip_context* axtbhip_context_create(int device, char* err, size_t errlen) {
  try {
    std::unique_ptr<axtbhip_context> c(new axtbhip_context);
    c->eng.reset(new Engine(device));
    c->eng_na.reset(new Engine(device));
    c->arena.base = (char*)scratch_address(device);
    c->arena.cap = NOALLOC_SCRATCH_BYTES;
    return c.release();
  } catch (std::exception& e) {
    if (err && errlen) snprintf(err, errlen, "%s", e.what());
    return nullptr;
  }
}

void axtbhip_context_destroy(axtbhip_context* ctx) { delete ctx; }

int axtbhip_last_error(axtbhip_context* ctx, char* buf, size_t len) {
  if (!ctx || !buf || !len) return 0;
  snprintf(buf, len, "%s", ctx->err.c_str());
  return (int)ctx->err.size();
}

int axtbhip_basis_size(const axtb_mol* mol, int* nsh, int* nao) {
  int s = 0, a = 0;
  for (int i = 0; i < mol->nat; i++) {
    int z = mol->numbers[i];
    if (z < 1 || z > 118) return AXTB_ERR_INPUT;
    if (z > AXTB_GFN2_MAX_ELEM) return AXTB_ERR_UNSUPPORTED_ELEMENT;
    for (int k = 0; k < GFN2_NSHELL[z - 1]; k++) {
      s++;
      a += 2 * GFN2_ANG_SHELL[z - 1][k] + 1;
    }
  }
  if (nsh) *nsh = s;
  if (nao) *nao = a;
  return 0;
}

int axtbhip_gfn2_engrad_batch_guess(axtbhip_context* ctx, const axtb_mol* mols, int nmol, const axtb_guess* guesses,
                                    const axtb_opts* opts, axtb_result* outs) {
  if (!ctx || !mols || !opts || !outs || nmol < 0) return AXTB_ERR_INPUT;
  std::vector<axtb::MolIn> in(nmol);
  for (int i = 0; i < nmol; i++) {
    in[i] = axtb::MolIn{mols[i].nat, mols[i].numbers, mols[i].positions, mols[i].charge, mols[i].uhf};
    if (guesses && guesses[i].shell_charges) {
      in[i].gq = guesses[i].shell_charges;
      in[i].gdp = guesses[i].dipoles;
      in[i].gqp = guesses[i].quadrupoles;
    }
  }
  std::vector<axtb::MolOut> res;
  try {
    run_warm_then_fresh(*ctx->eng, in, *opts, res);
  } catch (axtb::Fail& f) {
    ctx->err = f.msg;
    for (int i = 0; i < nmol; i++) outs[i].status = f.code;
    return f.code;
  } catch (std::exception& e) {
    ctx->err = e.what();
    for (int i = 0; i < nmol; i++) outs[i].status = AXTB_ERR_DEVICE;
    return AXTB_ERR_DEVICE;
  }
  int worst = 0;
  ctx->err.clear();
  for (int i = 0; i < nmol; i++) {
    fill_result(res[i], mols[i], &outs[i]);
    if (res[i].status && !worst) {
      worst = res[i].status;
      ctx->err = res[i].msg;
    }
  }
  return worst;
}

int axtbhip_gfn2_engrad_batch(axtbhip_context* ctx, const axtb_mol* mols, int nmol, const axtb_opts* opts,
                              axtb_result* outs) {
  return axtbhip_gfn2_engrad_batch_guess(ctx, mols, nmol, nullptr, opts, outs);
}

int axtbhip_gfn2_engrad(axtbhip_context* ctx, const axtb_mol* mol, const axtb_opts* opts, axtb_result* out) {
  return axtbhip_gfn2_engrad_batch(ctx, mol, 1, opts, out);
}

int axtbhip_gfn2_engrad_guess(axtbhip_context* ctx, const axtb_mol* mol, const axtb_opts* opts, const axtb_guess* guess,
                              axtb_result* out) {
  return axtbhip_gfn2_engrad_batch_guess(ctx, mol, 1, guess, opts, out);
}

int axtbhip_gfn2_hessian(axtbhip_context* ctx, const axtb_mol* mol, const axtb_opts* opts, const axtb_guess* guess,
                         double step, double accuracy, axtb_hessian* out) {
  if (!ctx || !mol || !opts || !out || mol->nat < 1 || !mol->numbers || !mol->positions) return AXTB_ERR_INPUT;
  out->status = AXTB_ERR_INPUT;
  if (!(step > 0.0 && std::isfinite(step))) {
    ctx->err = "the Hessian step must be a positive number";
    return AXTB_ERR_INPUT;
  }
  if (!(accuracy > 0.0 && std::isfinite(accuracy))) {
    ctx->err = "the Hessian's SCF accuracy must be a positive number";
    return AXTB_ERR_INPUT;
  }
  const int nat = mol->nat, n3 = 3 * nat;
  axtb_opts o = *opts;
  o.accuracy = std::min(accuracy, opts->accuracy);  // never looser than the calculator's
  o.want_gradient = 1;
  try {
    // the reference geometry, from the given start
    axtb::MolIn m0{nat, mol->numbers, mol->positions, mol->charge, mol->uhf};
    if (guess && guess->shell_charges) {
      m0.gq = guess->shell_charges;
      m0.gdp = guess->dipoles;
      m0.gqp = guess->quadrupoles;
    }
    std::vector<axtb::MolOut> r0;
    run_warm_then_fresh(*ctx->eng, {m0}, o, r0);
    if (r0[0].status) {
      ctx->err = r0[0].msg;
      out->status = r0[0].status;
      return r0[0].status;
    }
    // every coordinate forward by the step and back by twice the step (fd.rs:142-150, ForwardBack),
    // all 6 N displaced molecules in one batch, each started from the reference's moments
    std::vector<double> xyz((size_t)2 * n3 * n3);
    std::vector<axtb::MolIn> in(2 * n3);
    for (int y = 0; y < n3; y++)
      for (int sgn = 0; sgn < 2; sgn++) {
        double* x = &xyz[(size_t)(2 * y + sgn) * n3];
        std::copy(mol->positions, mol->positions + n3, x);
        x[y] = mol->positions[y] + step;
        if (sgn) x[y] = x[y] - 2.0 * step;
        in[2 * y + sgn] = axtb::MolIn{nat, mol->numbers, x, mol->charge, mol->uhf, r0[0].qsh.data(), r0[0].dp.data(),
                                      r0[0].qp.data()};
      }
    std::vector<axtb::MolOut> rd;
    run_warm_then_fresh(*ctx->eng, in, o, rd);
    for (int q = 0; q < 2 * n3; q++)
      if (rd[q].status) {  // the first failed displacement in coordinate order (fd.rs:170)
        ctx->err = "displacement " + std::to_string(q) + ": " + rd[q].msg;
        out->status = rd[q].status;
        return rd[q].status;
      }
    // column y = (g(x + h e_y) - g(x - h e_y)) / 2h, written as row y; then the symmetric part
    std::vector<double> H((size_t)n3 * n3);
    for (int y = 0; y < n3; y++) {
      const axtb::MolOut &p = rd[2 * y], &m = rd[2 * y + 1];
      for (int r = 0; r < n3; r++) H[(size_t)y * n3 + r] = (p.grad[r] - m.grad[r]) / (2.0 * step);
      if (out->dipole_derivatives)
        for (int k = 0; k < 3; k++) out->dipole_derivatives[3 * y + k] = (p.dipole[k] - m.dipole[k]) / (2.0 * step);
    }
    double asym = 0.0;  // fd.rs:128-140
    for (int i = 0; i < n3; i++)
      for (int j = 0; j < i; j++) {
        const double a = H[(size_t)i * n3 + j], b = H[(size_t)j * n3 + i];
        asym = std::max(asym, std::fabs(a - b));
        const double v = 0.5 * (a + b);
        H[(size_t)i * n3 + j] = H[(size_t)j * n3 + i] = v;
      }
    out->energy = r0[0].energy;
    if (out->gradient) std::memcpy(out->gradient, r0[0].grad.data(), sizeof(double) * n3);
    if (out->hessian) std::memcpy(out->hessian, H.data(), sizeof(double) * H.size());
    out->max_asymmetry = asym;
    out->displacements = 2 * n3;
    out->status = 0;
    ctx->err.clear();
    return 0;
  } catch (axtb::Fail& f) {
    ctx->err = f.msg;
    out->status = f.code;
    return f.code;
  } catch (std::exception& e) {
    ctx->err = e.what();
    out->status = AXTB_ERR_DEVICE;
    return AXTB_ERR_DEVICE;
  }
}

int axtbhip_gfn2_engrad_noalloc(axtbhip_context* ctx, const axtb_mol* mol, const axtb_opts* opts, axtb_result* out) {
  if (!ctx || !mol || !opts || !out || mol->nat < 0 || (mol->nat > 0 && (!mol->numbers || !mol->positions)))
    return AXTB_ERR_INPUT;
  int nsh = 0, nao = 0;
  if (axtbhip_basis_size(mol, &nsh, &nao) == 0 && nao > NOALLOC_MAX_NAO) {
    ctx->err = "the no-allocation entry point takes at most " + std::to_string(NOALLOC_MAX_NAO) +
               " basis functions (this molecule has " + std::to_string(nao) + ")";
    out->status = AXTB_ERR_INPUT;
    return AXTB_ERR_INPUT;
  }
  std::vector<axtb::MolIn> in{axtb::MolIn{mol->nat, mol->numbers, mol->positions, mol->charge, mol->uhf}};
  std::vector<axtb::MolOut> res;
  // every buffer of this call is carved from the device's scratch; nothing is allocated on the
  // device or pinned on the host
  struct Scope {
    axtb::Arena& a;
    explicit Scope(axtb::Arena& x) : a(x) {
      a.reset();
      axtb::t_arena = &a;
    }
    ~Scope() { axtb::t_arena = nullptr; }
  } scope(ctx->arena);
  try {
    ctx->eng_na->prof.on = false;
    ctx->eng_na->run(in, *opts, res);
  } catch (axtb::Fail& f) {
    ctx->err = f.msg;
    out->status = f.code;
    return f.code;
  } catch (std::exception& e) {
    ctx->err = e.what();
    out->status = AXTB_ERR_DEVICE;
    return AXTB_ERR_DEVICE;
  }
  fill_result(res[0], *mol, out);
  ctx->err = res[0].status ? res[0].msg : std::string();
  return res[0].status;
}

int axtbhip_batch_prepare_guess(axtbhip_context* ctx, const axtb_mol* mols, int nmol, const axtb_guess* guesses,
                                const axtb_opts* opts) {
  if (!ctx || !mols || !opts || nmol < 0) return AXTB_ERR_INPUT;
  std::vector<axtb::MolIn> in(nmol);
  for (int i = 0; i < nmol; i++) {
    in[i] = axtb::MolIn{mols[i].nat, mols[i].numbers, mols[i].positions, mols[i].charge, mols[i].uhf};
    if (guesses && guesses[i].shell_charges) {
      in[i].gq = guesses[i].shell_charges;
      in[i].gdp = guesses[i].dipoles;
      in[i].gqp = guesses[i].quadrupoles;
    }
  }
  try {
    ctx->eng->prepare_resident(in, *opts);
    ctx->err.clear();
    return 0;
  } catch (axtb::Fail& f) {
    ctx->err = f.msg;
    return f.code;
  } catch (std::exception& e) {
    ctx->err = e.what();
    return AXTB_ERR_DEVICE;
  }
}

int axtbhip_batch_prepare(axtbhip_context* ctx, const axtb_mol* mols, int nmol, const axtb_opts* opts) {
  return axtbhip_batch_prepare_guess(ctx, mols, nmol, nullptr, opts);
}

int axtbhip_batch_update(axtbhip_context* ctx, const axtb_mol* mols, int nmol, const axtb_guess* guesses) {
  if (!ctx || !mols || nmol < 0) return AXTB_ERR_INPUT;
  if (!ctx->eng->resident || nmol != (int)ctx->eng->prep_out.size()) {
    ctx->err = "no prepared batch of this size: batch_prepare first";
    return AXTB_ERR_INPUT;
  }
  std::vector<axtb::MolIn> in(nmol);
  for (int i = 0; i < nmol; i++) {
    in[i] = axtb::MolIn{mols[i].nat, mols[i].numbers, mols[i].positions, mols[i].charge, mols[i].uhf};
    if (guesses && guesses[i].shell_charges) {
      in[i].gq = guesses[i].shell_charges;
      in[i].gdp = guesses[i].dipoles;
      in[i].gqp = guesses[i].quadrupoles;
    }
  }
  try {
    ctx->eng->update(in);
    ctx->err.clear();
    return 0;
  } catch (axtb::Fail& f) {
    ctx->err = f.msg;
    return f.code;
  } catch (std::exception& e) {
    ctx->err = e.what();
    return AXTB_ERR_DEVICE;
  }
}

int axtbhip_batch_run(axtbhip_context* ctx) {
  if (!ctx) return AXTB_ERR_INPUT;
  if (!ctx->eng->resident) {
    ctx->err = "no prepared batch: batch_prepare first";
    return AXTB_ERR_INPUT;
  }
  try {
    ctx->eng->compute();
    return 0;
  } catch (axtb::Fail& f) {
    ctx->err = f.msg;
    return f.code;
  } catch (std::exception& e) {
    ctx->err = e.what();
    return AXTB_ERR_DEVICE;
  }
}

int axtbhip_batch_gradient(axtbhip_context* ctx) {
  if (!ctx) return AXTB_ERR_INPUT;
  if (!ctx->eng->resident) {
    ctx->err = "no prepared batch: batch_prepare first";
    return AXTB_ERR_INPUT;
  }
  try {
    ctx->eng->gradient_of_state();
    return 0;
  } catch (axtb::Fail& f) {
    ctx->err = f.msg;
    return f.code;
  } catch (std::exception& e) {
    ctx->err = e.what();
    return AXTB_ERR_DEVICE;
  }
}

int axtbhip_batch_results(axtbhip_context* ctx, const axtb_mol* mols, int nmol, axtb_result* outs) {
  if (!ctx || !outs || nmol < 0) return AXTB_ERR_INPUT;
  Engine& E = *ctx->eng;
  if (!E.resident || nmol != (int)E.prep_out.size()) {
    ctx->err = "no prepared batch of this size: batch_prepare first";
    return AXTB_ERR_INPUT;
  }
  if (!E.scf_done && !E.mols.empty()) {
    ctx->err = "the prepared batch has no results: batch_run first";
    return AXTB_ERR_INPUT;
  }
  try {
    std::vector<axtb::MolOut> res = E.prep_out;
    if (!E.mols.empty()) E.collect(E.opts_, res);
    int worst = 0;
    for (int i = 0; i < nmol; i++) {
      fill_result(res[i], mols[i], &outs[i]);
      if (res[i].status && !worst) {
        worst = res[i].status;
        ctx->err = res[i].msg;
      }
    }
    return worst;
  } catch (std::exception& e) {
    ctx->err = e.what();
    return AXTB_ERR_DEVICE;
  }
}

int axtbhip_batch_adjoint(axtbhip_context* ctx, const double* dE, double* grad) {
  if (!ctx || !dE || !grad) return AXTB_ERR_INPUT;
  Engine& E = *ctx->eng;
  if (!E.resident) {
    ctx->err = "no prepared batch: batch_prepare first";
    return AXTB_ERR_INPUT;
  }
  try {
    E.gradient_of_state();
    const int nin = (int)E.prep_out.size();
    // input offsets of the atoms; inputs that failed at prepare hold zeros
    std::vector<long long> off(nin + 1, 0);
    std::vector<int> nat_in(nin, 0);
    for (int m = 0; m < E.nmol(); m++) nat_in[E.map_use[m]] = E.mols[m].nat;
    for (int i = 0; i < nin; i++) off[i + 1] = off[i] + nat_in[i];
    std::vector<axtb::MolOut> res = E.prep_out;
    if (!E.mols.empty()) E.collect(E.opts_, res);
    std::memset(grad, 0, sizeof(double) * 3 * off[nin]);
    for (int i = 0; i < nin; i++)
      if (res[i].status == 0 && (int)res[i].grad.size() == 3 * nat_in[i])
        for (int k = 0; k < 3 * nat_in[i]; k++) grad[3 * off[i] + k] = dE[i] * res[i].grad[k];
    return 0;
  } catch (axtb::Fail& f) {
    ctx->err = f.msg;
    return f.code;
  } catch (std::exception& e) {
    ctx->err = e.what();
    return AXTB_ERR_DEVICE;
  }
}

int axtbhip_tuning_set(axtbhip_context* ctx, const char* knob, int variant, int bucket, int value) {
  if (!ctx || !knob || bucket < 0 || bucket >= axtb::Engine::TUNE_NB) return AXTB_ERR_INPUT;
  axtb::Engine& E = *ctx->eng;
  const std::string k(knob);
  E.list_sig.clear();  // the fixed eigensolver plans on the device hold launch shapes: build them again
  if (k == "jacobi_threads" && variant >= 0 && variant < 6 && (value == 0 || value == 256 || value == 512 ||
                                                                 value == 1024)) {
    E.tune_jac_threads[variant / 3][variant % 3][bucket] = value;
    return 0;
  }
  if (k == "jacobi_mode" && (variant == 0 || variant == 1) && value >= 0 && value <= 2) {
    E.tune_jac_mode[variant][bucket] = value;
    return 0;
  }
  if (k == "gemm_small" && variant == 0 && value >= 0 && value <= 2) {
    E.tune_gemm_small[bucket] = value;
    return 0;
  }
  ctx->err = "unknown tuning knob or value";
  return AXTB_ERR_INPUT;
}

int axtbhip_synchronize(axtbhip_context* ctx) {
  if (!ctx) return AXTB_ERR_INPUT;
  return hipDeviceSynchronize() == hipSuccess ? 0 : AXTB_ERR_DEVICE;
}

void axtbhip_profile_enable(axtbhip_context* ctx, int on) {
  if (ctx) {
    ctx->eng->prof.on = on != 0;
    if (!on) ctx->eng->prof.drain();
  }
}

int axtbhip_profile_report(axtbhip_context* ctx, char* buf, size_t len) {
  if (!ctx || !buf || !len) return 0;
  try {
    std::string out;
    ctx->eng->drain_sweeps();
    for (auto& kv : ctx->eng->prof.drain()) {
      char line[256];
      snprintf(line, sizeof line, "%s %ld %.6f %.6e %.6f\n", kv.first.c_str(), kv.second.n, kv.second.ms,
               kv.second.work, kv.second.span);
      out += line;
    }
    snprintf(buf, len, "%s", out.c_str());
    return (int)out.size();
  } catch (std::exception&) {
    return 0;
  }
}

}  // extern "C"

// ================================================================================ test hooks
// Kernel-level entry points for the parity tests (declared in include/axtb_hip_test.h): each
// runs one piece of the engine on host arrays.
extern "C" {

int axtbhip_test_gemm(int m, int n, int k, int ta, int tb, const double* A, int lda, const double* B, int ldb,
                      const double* dk, double* C, int ldc) {
  try {
    using namespace axtb;
    Engine e(0);
    DBuf<double> a, b, c, d;
    size_t na = (size_t)lda * (ta ? m : k), nb = (size_t)ldb * (tb ? k : n), nc = (size_t)ldc * n;
    a.upload(std::vector<double>(A, A + na), e.st);
    b.upload(std::vector<double>(B, B + nb), e.st);
    c.ensure(nc);
    if (dk) d.upload(std::vector<double>(dk, dk + k), e.st);
    std::vector<GemmJob> j{GemmJob{m, n, k, ta, tb, a.p, b.p, c.p, dk ? d.p : nullptr, lda, ldb, ldc, 1.0, 0.0}};
    e.gemm(j);
    HIP_OK(hipMemcpyAsync(C, c.p, nc * sizeof(double), hipMemcpyDeviceToHost, e.st));
    HIP_OK(hipStreamSynchronize(e.st));
    return 0;
  } catch (std::exception&) {
    return AXTB_ERR_DEVICE;
  }
}

int axtbhip_test_gemm_upd32(int m, int n, const double* U, int ldu, const double* B, int ldb, double* C, int ldc) {
  try {
    using namespace axtb;
    Engine e(0);
    DBuf<double> u, b, c;
    u.upload(std::vector<double>(U, U + (size_t)ldu * n), e.st);
    b.upload(std::vector<double>(B, B + (size_t)ldb * n), e.st);
    c.ensure((size_t)ldc * n);
    std::vector<GemmJob> j{GemmJob{m, n, n, 0, 0, u.p, b.p, c.p, nullptr, ldu, ldb, ldc, 1.0, 0.0}};
    j[0].f32upd = 1;
    e.gemm(j);
    HIP_OK(hipMemcpyAsync(C, c.p, (size_t)ldc * n * sizeof(double), hipMemcpyDeviceToHost, e.st));
    HIP_OK(hipStreamSynchronize(e.st));
    return 0;
  } catch (std::exception&) {
    return AXTB_ERR_DEVICE;
  }
}

// eigenpairs of `nb` symmetric n x n matrices (column-major, consecutive): w [nb][n], Z [nb][n][n]
int axtbhip_test_eigh(int n, int nb, const double* A, double* w, double* Z, int* sweeps) {
  try {
    using namespace axtb;
    Engine e(0);
    DBuf<double> a, v, ww, z;
    size_t nn = (size_t)n * n;
    a.upload(std::vector<double>(A, A + nn * nb), e.st);
    v.ensure(nn * nb);
    ww.ensure((size_t)n * nb);
    z.ensure(nn * nb);
    std::vector<EigJob> jobs(nb);
    for (int i = 0; i < nb; i++) jobs[i] = EigJob{n, a.p + nn * i, v.p + nn * i, ww.p + (size_t)n * i, z.p + nn * i};
    e.eigh(jobs);
    HIP_OK(hipMemcpyAsync(w, ww.p, (size_t)n * nb * sizeof(double), hipMemcpyDeviceToHost, e.st));
    HIP_OK(hipMemcpyAsync(Z, z.p, nn * nb * sizeof(double), hipMemcpyDeviceToHost, e.st));
    if (sweeps) HIP_OK(hipMemcpyAsync(sweeps, e.eadhoc.stat.p, nb * sizeof(int), hipMemcpyDeviceToHost, e.st));
    HIP_OK(hipStreamSynchronize(e.st));
    return 0;
  } catch (std::exception&) {
    return AXTB_ERR_DEVICE;
  }
}

// The dense eigensolver (tridiagonalisation, divide and conquer, back-transformation) on one
// symmetric n x n matrix: all eigenvalues w, the first k eigenvectors Z (n x k).
int axtbhip_test_dense_eigh(int n, const double* A, int k, double* w, double* Z) {
  try {
    using namespace axtb;
    Engine e(0);
    DBuf<double> a, ww, z, tau, q;
    a.upload(std::vector<double>(A, A + (size_t)n * n), e.st);
    ww.ensure(n);
    tau.ensure(n);
    q.ensure((size_t)n * n);
    z.ensure((size_t)n * std::max(k, 1));
    e.dense_values(a.p, n, ww.p, tau.p, q.p);
    e.dense_vectors(a.p, n, tau.p, q.p, k, z.p);
    HIP_OK(hipMemcpyAsync(w, ww.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, e.st));
    if (k > 0) HIP_OK(hipMemcpyAsync(Z, z.p, (size_t)n * k * sizeof(double), hipMemcpyDeviceToHost, e.st));
    HIP_OK(hipStreamSynchronize(e.st));
    return 0;
  } catch (std::exception&) {
    return AXTB_ERR_DEVICE;
  }
}

// Setup of one molecule: S, H0 (nao x nao column-major), D (3 nao^2), Q (6 nao^2), the GFN2 CN,
// the start shell charges, X = S^-1/2 and the D4 three-body energy per atom.
int axtbhip_test_setup(const axtb_mol* mol, const axtb_opts* opts, double* S, double* H0, double* D, double* Q,
                       double* cn, double* qstart, double* X, double* e3) {
  try {
    using namespace axtb;
    Engine e(0);
    std::vector<MolIn> in{MolIn{mol->nat, mol->numbers, mol->positions, mol->charge, mol->uhf}};
    std::vector<MolOut> out(1);
    e.build(in, {0}, *opts, out);
    if (e.mols.empty()) return out[0].status ? out[0].status : AXTB_ERR_INPUT;
    e.setup(*opts);
    size_t nn = (size_t)e.mols[0].nao * e.mols[0].nao;
    auto cp = [&](double* dst, const double* src, size_t n) {
      if (dst) HIP_OK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, e.st));
    };
    cp(S, e.pool.p, nn);
    cp(H0, e.pool.p + e.tot_mat, nn);
    for (int k = 0; k < 3; k++) cp(D ? D + k * nn : nullptr, e.pool.p + (2 + k) * e.tot_mat, nn);
    for (int k = 0; k < 6; k++) cp(Q ? Q + k * nn : nullptr, e.pool.p + (5 + k) * e.tot_mat, nn);
    cp(cn, e.cn.p, mol->nat);
    cp(qstart, e.qstart.p, e.mols[0].nsh);
    cp(X, e.X.p, nn);
    cp(e3, e.e3.p, mol->nat);
    HIP_OK(hipStreamSynchronize(e.st));
    return 0;
  } catch (axtb::Fail& f) {
    return f.code;
  } catch (std::exception&) {
    return AXTB_ERR_DEVICE;
  }
}

// The first SCC iteration of one molecule from its start: F, eigenvalues, P and the output
// shell charges, plus the potentials it used (vsh [nsh], vat [nat], vdp [3 nat], vqp [6 nat]).
int axtbhip_test_iter1(const axtb_mol* mol, const axtb_opts* opts, double* F, double* w, double* P, double* qsh_out,
                       double* vsh, double* vat, double* vdp, double* vqp) {
  try {
    using namespace axtb;
    Engine e(0);
    std::vector<MolIn> in{MolIn{mol->nat, mol->numbers, mol->positions, mol->charge, mol->uhf}};
    std::vector<MolOut> out(1);
    e.build(in, {0}, *opts, out);
    if (e.mols.empty()) return out[0].status ? out[0].status : AXTB_ERR_INPUT;
    e.setup(*opts);
    axtb_opts o1 = *opts;
    o1.max_iter = 1;
    e.scf(o1);
    size_t nn = (size_t)e.mols[0].nao * e.mols[0].nao;
    auto cp = [&](double* dst, const double* src, size_t n) {
      if (dst) HIP_OK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, e.st));
    };
    cp(F, e.F.p, nn);
    cp(w, e.w.p, e.mols[0].nao);
    cp(P, e.P.p, nn);
    cp(qsh_out, e.qsh_o.p, e.mols[0].nsh);
    cp(vsh, e.vsh.p, e.mols[0].nsh);
    cp(vat, e.vat.p, mol->nat);
    cp(vdp, e.vdp.p, 3 * mol->nat);
    cp(vqp, e.vqp.p, 6 * mol->nat);
    HIP_OK(hipStreamSynchronize(e.st));
    return 0;
  } catch (axtb::Fail& f) {
    return f.code;
  } catch (std::exception&) {
    return AXTB_ERR_DEVICE;
  }
}

}  // extern "C"
