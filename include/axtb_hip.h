/* axtb-hip: GFN2-xTB energy + analytic gradient on AMD MI355X (gfx950), C ABI.
 *
 * The drop-in boundary of axtb's GFN2 calculator. Each entry point replaces one call of the
 * reference's Rust API (paths relative to the axtb checkout):
 *
 *   axtbhip_gfn2_engrad        Method::singlepoint(positions, Want::GRADIENT)
 *                              (crates/base/src/method/mod.rs:62, implemented for GFN2 by
 *                              axtb::Calculator, crates/axtb/src/calculator.rs:378-478, which runs
 *                              XtbModel::singlepoint_state, crates/axtb/src/xtb.rs:437-861)
 *   axtbhip_gfn2_engrad_batch  many independent Method::singlepoint calls, as
 *                              Session.compute_many does (crates/axtb-py/src/engine.rs:844)
 *   axtbhip_basis_size         the AO/shell counts of axtb::basis::Basis (crates/axtb/src/basis.rs:165-229)
 *
 * Conventions: positions in bohr, row-major [nat][3]; energies in Hartree, gradients in
 * Hartree/bohr; axtb's default numerical conventions (crates/axtb/src/conventions.rs:244-268).
 * All arrays are caller-owned; result pointers may be NULL to skip an output.
 * Status codes mirror axtb_base::Error (crates/base/src/error.rs:8-24).
 */
#ifndef AXTB_HIP_H
#define AXTB_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AXTB_OK 0
#define AXTB_ERR_INPUT 1               /* Error::Input */
#define AXTB_ERR_NUMERICAL 2           /* Error::Numerical */
#define AXTB_ERR_SCF_NOT_CONVERGED 3   /* Error::ScfNotConverged */
#define AXTB_ERR_UNSUPPORTED_ELEMENT 4 /* Error::UnsupportedElement */
#define AXTB_ERR_DEVICE 5              /* a HIP runtime failure (no counterpart in the reference) */

/* A molecule: Molecule::new(numbers, positions, charge, uhf) (crates/base/src/molecule.rs:35-64). */
typedef struct axtb_mol {
  int nat;
  const int* numbers;      /* atomic numbers [nat] */
  const double* positions; /* bohr [nat][3] */
  double charge;           /* molecular charge (integer valued) */
  int uhf;                 /* unpaired electrons; -1: the lowest spin of the electron count */
} axtb_mol;

/* Options of one calculation: axtb::Options / ScfOptions (crates/axtb/src/lib.rs:92-153,
 * crates/axtb/src/xtb.rs:65-99). */
#define AXTB_PREC_FP64 0
#define AXTB_PREC_MIXED 1
#define AXTB_PREC_FP64_REFINE 2

typedef struct axtb_opts {
  double accuracy;   /* SCF accuracy: energy threshold 1e-6*acc, residual 2e-5*acc (default 1) */
  double etemp;      /* electronic temperature in K (default 300) */
  int max_iter;      /* SCF iterations (default 250) */
  double damping;    /* Broyden damping (default 0.4) */
  int guess;         /* 0: EEQ charges (default), 1: neutral atoms */
  int want_gradient; /* 1: analytic gradient */
  /* AXTB_PREC_FP64 (0): FP64 throughout, a full eigensolve every SCC iteration; AXTB_PREC_MIXED (1):
   * FP32 Jacobi and FP32 refinement updates, FP64 Ogita-Aishima refinement, later iterations refine
   * the previous orbitals; AXTB_PREC_FP64_REFINE (2): the refinement of mode 1 with every step in
   * FP64 (no counterpart in the reference, which solves in full every iteration) */
  int precision;
  int dispersion;    /* 1: D4 (default) */
  /* SCF mixer plan (MixerKind, crates/axtb/src/scf/mixing.rs:83-118): 0 the conventions' default
   * (xtb.rs:516; axtb, conventions.rs:260: Broyden, then DIIS from the start if Broyden does not
   * converge; tblite: Broyden), 1 Broyden alone, 2 DIIS alone, 3 Broyden then DIIS under any
   * conventions */
  int mixer;
  /* implicit solvation (SolvationSpec, crates/axtb/src/solvation.rs:14-22; the axtb conventions'
   * corrected ALPB/GBSA model, conventions.rs:256): AXTB_SOLV_NONE, AXTB_SOLV_ALPB or
   * AXTB_SOLV_GBSA; the solvent's name as xtb accepts it (any case, xtb's aliases; copied by the
   * call); the reference state AXTB_SOLV_GSOLV, AXTB_SOLV_BAR1M or AXTB_SOLV_REFERENCE. An unknown
   * solvent or one without parameters for the model is an AXTB_ERR_INPUT of the call. */
  int solvation;
  const char* solvent;
  int solvation_state;
  /* numerical conventions (Conventions, crates/axtb/src/conventions.rs:73-84, 237-353):
   * AXTB_CONV_AXTB (default) or AXTB_CONV_TBLITE (tblite 0.8.0: its kB, mctc-lib's bohr in the radius
   * tables, 25 bohr CN and repulsion cutoffs, D4 cutoffs 30/50/25 bohr with 0.05 bohr switches,
   * primitive screening at accuracy 1 and atom pairs screened by the basis extent, zero-charge
   * start, Broyden mixing); the xtb programs' conventions are not run by the GPU engine
   * (AXTB_ERR_INPUT). */
  int conventions;
} axtb_opts;

#define AXTB_CONV_AXTB 0
#define AXTB_CONV_XTB 1
#define AXTB_CONV_XTB671 2
#define AXTB_CONV_TBLITE 3

#define AXTB_SOLV_NONE 0
#define AXTB_SOLV_ALPB 1
#define AXTB_SOLV_GBSA 2
#define AXTB_SOLV_GSOLV 0
#define AXTB_SOLV_BAR1M 1
#define AXTB_SOLV_REFERENCE 2

#define AXTB_MIXER_AUTO 0
#define AXTB_MIXER_BROYDEN 1
#define AXTB_MIXER_DIIS 2
#define AXTB_MIXER_BROYDEN_DIIS 3

/* Results: the SinglePoint fields of the path (crates/base/src/method/results.rs:39-230). */
typedef struct axtb_result {
  double energy;
  double* gradient;         /* [nat][3] or NULL */
  double* charges;          /* [nat] or NULL */
  double* shell_charges;    /* [nsh] or NULL */
  double* dipoles;          /* atomic dipoles [nat][3] or NULL */
  double* quadrupoles;      /* atomic traceless quadrupoles [nat][6] (xx xy yy xz yz zz) or NULL */
  double* orbital_energies; /* [nao] or NULL */
  /* repulsion, dispersion, electronic, isotropic_es, third_order, anisotropic, entropy, band,
   * solvation (EnergyTerms, crates/axtb/src/xtb.rs:109-125: the solvation free energy; its
   * charge-dependent part is also part of electronic) */
  double terms[9];
  double efermi, residual, energy_change;
  int iterations, status, nsh, nao;
  /* SinglePoint.dipole: molecular dipole about the origin, e bohr (crates/axtb/src/xtb.rs:746-751) */
  double dipole[3];
  /* HOMO-LUMO gap above the alpha HOMO in Eh (properties.rs:204-240); NaN without a LUMO */
  double gap;
  /* ScfReport (crates/base/src/method/results.rs:156-176): the mixer of the result (0 broyden,
   * 1 diis, 2 diis-fallback), the attempts of the plan and each one's iterations and outcome.
   * iterations above is the converged attempt's; on AXTB_ERR_SCF_NOT_CONVERGED, all attempts'. */
  int mixer_used, nattempts, attempt_iterations[2], attempt_converged[2];
  /* AXTB_ERR_UNSUPPORTED_ELEMENT: the element (Error::UnsupportedElement{z}, error.rs:12), else 0 */
  int element;
} axtb_result;

typedef struct axtbhip_context axtbhip_context;

const char* axtbhip_version(void);
/* One context per host thread and device (it owns a HIP stream and its device buffers). */
axtbhip_context* axtbhip_context_create(int device, char* err, size_t errlen);
void axtbhip_context_destroy(axtbhip_context* ctx);
int axtbhip_last_error(axtbhip_context* ctx, char* buf, size_t len);
int axtbhip_basis_size(const axtb_mol* mol, int* nsh, int* nao);
int axtbhip_gfn2_engrad(axtbhip_context* ctx, const axtb_mol* mol, const axtb_opts* opts, axtb_result* out);
/* Returns 0 if every molecule succeeded, else the status of the first that failed; each
 * out[i].status holds the molecule's own status. */
int axtbhip_gfn2_engrad_batch(axtbhip_context* ctx, const axtb_mol* mols, int nmol, const axtb_opts* opts,
                              axtb_result* outs);

/* A given SCF start of one molecule: the converged moments of a previous geometry, as a
 * Calculator keeps them between single points (ScfGuess, crates/axtb/src/xtb.rs:204-215; used
 * at xtb.rs:519-529). Array sizes are those of the molecule (nsh from axtbhip_basis_size) and
 * their layouts those of axtb_result. */
typedef struct axtb_guess {
  const double* shell_charges; /* [nsh]; NULL: no given start, opts->guess decides */
  const double* dipoles;       /* [nat][3]; NULL: zero */
  const double* quadrupoles;   /* [nat][6] (xx xy yy xz yz zz); NULL: zero */
} axtb_guess;

/* axtbhip_gfn2_engrad_batch with a given start per molecule (guesses may be NULL: none). A
 * molecule that started from its given moments and did not converge is computed once more from
 * the start opts->guess names, as Calculator::singlepoint_inner does with
 * RetryLadder::warm_then_fresh (crates/axtb/src/calculator.rs:194-206,
 * crates/base/src/method/retry.rs:150-158). */
int axtbhip_gfn2_engrad_batch_guess(axtbhip_context* ctx, const axtb_mol* mols, int nmol, const axtb_guess* guesses,
                                    const axtb_opts* opts, axtb_result* outs);

/* axtbhip_gfn2_engrad with a given SCF start (guess may be NULL: the start opts->guess names):
 * Method::singlepoint on a calculator holding a wavefunction (set_wavefunction,
 * crates/base/src/method/mod.rs:71-101; crates/axtb/src/calculator.rs:185-206), including its
 * retry from the fresh start when the given start does not converge. */
int axtbhip_gfn2_engrad_guess(axtbhip_context* ctx, const axtb_mol* mol, const axtb_opts* opts,
                              const axtb_guess* guess, axtb_result* out);

/* A semi-numerical Hessian (HessianMethod::Numerical of axtb's GFN2 calculator,
 * crates/axtb/src/calculator.rs:266-296, over crates/base/src/fd.rs:163-209): the reference
 * geometry's energy and gradient (from `guess` if given), then every Cartesian coordinate moved
 * forward by `step` bohr and back by twice the step, the 6 N displaced gradients computed as one
 * GPU batch, each SCF started from the reference's moments (a displacement that does not converge
 * from them runs once more from the fresh start) at SCF accuracy min(accuracy, opts->accuracy).
 * Column y of the Hessian is the central difference of the gradients, the matrix is symmetrized
 * (fd.rs:128-140). The reference's defaults: step 0.005 bohr, accuracy 0.01
 * (calculator.rs:81-86). A non-positive or non-finite step or accuracy is AXTB_ERR_INPUT; a failed
 * displacement returns its status (the first in coordinate order). */
#define AXTB_HESSIAN_STEP 0.005
#define AXTB_HESSIAN_ACCURACY 0.01
typedef struct axtb_hessian {
  double energy;               /* at the reference geometry */
  double* gradient;            /* [nat][3] or NULL */
  double* hessian;             /* [3 nat][3 nat] row-major, Eh/bohr^2, or NULL */
  double* dipole_derivatives;  /* d mu / d R_x, [3 nat][3] (e), or NULL */
  double max_asymmetry;        /* largest |H_ij - H_ji| before the symmetrization */
  int displacements;           /* displaced single points computed (6 nat) */
  int status;
} axtb_hessian;
int axtbhip_gfn2_hessian(axtbhip_context* ctx, const axtb_mol* mol, const axtb_opts* opts, const axtb_guess* guess,
                         double step, double accuracy, axtb_hessian* out);

/* axtbhip_gfn2_engrad that allocates nothing: no counterpart in the reference. For callers that
 * cannot allocate device or pinned host memory during the call (code under HIP graph capture,
 * tight serving loops, harnesses that forbid it). Every device buffer of the calculation is carved
 * from one static 64 MiB array per GPU (held until the process ends); molecules of at most 256
 * basis functions (AXTB_ERR_INPUT beyond, and if a calculation would need more of the array).
 * All calls on one GPU share that array and the last call's contents stay in it: calls that
 * overlap on one GPU (from several host threads or contexts) race. */
int axtbhip_gfn2_engrad_noalloc(axtbhip_context* ctx, const axtb_mol* mol, const axtb_opts* opts, axtb_result* out);

/* A batch kept on the device between calls, for repeated or timed runs: prepare validates and
 * uploads the molecules (host arrays may be released afterwards), run computes on the resident
 * inputs, results downloads them (mols gives the array sizes). No counterpart in the reference:
 * its calculators own their inputs the same way (crates/axtb/src/calculator.rs:31-52). */
int axtbhip_batch_prepare(axtbhip_context* ctx, const axtb_mol* mols, int nmol, const axtb_opts* opts);
/* batch_prepare with a given start per molecule (guesses may be NULL), copied with the batch: every
 * following batch_run starts those molecules from them (a trajectory step of many molecules with
 * its inputs resident). No retry here: a warm start that does not converge reports
 * AXTB_ERR_SCF_NOT_CONVERGED, and the caller runs it again without a start. */
int axtbhip_batch_prepare_guess(axtbhip_context* ctx, const axtb_mol* mols, int nmol, const axtb_guess* guesses,
                                const axtb_opts* opts);
/* New positions (and starts; guesses may be NULL: none) for the molecules of the prepared batch, the
 * next step of a trajectory without a new prepare: mols must hold the prepared molecules' atoms in
 * the same order (nmol as at prepare; molecules that failed there are skipped). Checked as at
 * prepare before anything changes; any mismatch returns AXTB_ERR_INPUT and keeps the batch. The
 * counterpart of Method::singlepoint(new positions) on a calculator that keeps its molecule
 * (crates/axtb/src/calculator.rs:194-206). */
int axtbhip_batch_update(axtbhip_context* ctx, const axtb_mol* mols, int nmol, const axtb_guess* guesses);
int axtbhip_batch_run(axtbhip_context* ctx);
int axtbhip_batch_results(axtbhip_context* ctx, const axtb_mol* mols, int nmol, axtb_result* outs);
/* The adjoint step on the resident batch: after a batch_run with want_gradient = 0, the analytic
 * gradient of the converged state (W and P contracted with dS/dR and dH0/dR recomputed per
 * shell-pair block, plus the ES2/ES3/AES/D4/repulsion and CN-chain terms), equal to a run with
 * want_gradient = 1; batch_results then returns it. Replaces XtbModel::gradient_functional
 * (crates/axtb/src/xtb.rs:392-421) applied to the state of the last single point. Returns
 * AXTB_ERR_INPUT before any batch_run. */
int axtbhip_batch_gradient(axtbhip_context* ctx);
/* The vector-Jacobian product of the resident batch's energies (the backward of an energy
 * function, XtbModel::gradient_functional, crates/axtb/src/xtb.rs:392-421, scaled): grad receives
 * dE[i] * dE_i/dR for every prepared molecule i, packed in input order ([sum of nat_i][3], bohr);
 * molecules that failed get zeros. Runs batch_gradient first if needed. */
int axtbhip_batch_adjoint(axtbhip_context* ctx, const double* dE, double* grad);

/* Per-kernel timing with HIP events on the context's stream (the counterpart of the reference's
 * phase timers, crates/base/src/timing.rs:11-54): report lines "kernel launches total_ms work
 * span_ms", work = the launches' algorithmic flops (GEMM, eigensolver) or HBM bytes (streaming
 * kernels), span_ms = the time covered by the union of the launches' intervals (launches on two
 * streams overlap, so span_ms can be below total_ms). */
void axtbhip_profile_enable(axtbhip_context* ctx, int on);
/* hipDeviceSynchronize on the context's device. */
int axtbhip_synchronize(axtbhip_context* ctx);
/* Launch shapes per size bucket (n / 8, buckets 0..25, n up to 207) as an autotuner measured them
 * (axtb_amd/autotune.py; no counterpart in the reference, which has no GPU path): knob
 * "jacobi_threads" (variant 3 * fp32 + mode, mode 1: A in LDS, 2: A and V in LDS; value 256, 512
 * or 1024 threads per matrix, applied where the matrix's 2 x 2 blocks fit), "jacobi_mode" (variant
 * fp32; value 1: keep the eigenvectors in HBM where both fit in LDS) and "gemm_small" (variant 0;
 * value 1: products up to 128 x 128 on the one-workgroup kernel, 2: on the 64-tile kernel).
 * Value 0 restores the built-in rule. Returns 0 or AXTB_ERR_INPUT. */
int axtbhip_tuning_set(axtbhip_context* ctx, const char* knob, int variant, int bucket, int value);
int axtbhip_profile_report(axtbhip_context* ctx, char* buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
