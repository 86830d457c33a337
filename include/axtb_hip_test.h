/* Kernel-level test hooks of libaxtb_hip.so (parity tests only; not part of the drop-in API). */
#ifndef AXTB_HIP_TEST_H
#define AXTB_HIP_TEST_H
#include "axtb_hip.h"
#ifdef __cplusplus
extern "C" {
#endif
int axtbhip_test_gemm(int m, int n, int k, int ta, int tb, const double* A, int lda, const double* B, int ldb,
                      const double* dk, double* C, int ldc);
/* C = U + U (B - I) (U m x n, B n x n) with the product on the FP32 matrix cores */
int axtbhip_test_gemm_upd32(int m, int n, const double* U, int ldu, const double* B, int ldb, double* C, int ldc);
int axtbhip_test_eigh(int n, int nb, const double* A, double* w, double* Z, int* sweeps);
int axtbhip_test_dense_eigh(int n, const double* A, int k, double* w, double* Z);
int axtbhip_test_setup(const axtb_mol* mol, const axtb_opts* opts, double* S, double* H0, double* D, double* Q,
                       double* cn, double* qstart, double* X, double* e3);
int axtbhip_test_iter1(const axtb_mol* mol, const axtb_opts* opts, double* F, double* w, double* P, double* qsh_out,
                       double* vsh, double* vat, double* vdp, double* vqp);
#ifdef __cplusplus
}
#endif
#endif
