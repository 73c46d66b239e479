/*
 * oracle.h -- FP64 CPU restatement of MM_Propagation's Lawson-Runge-Kutta
 * CPU path (src/MultimodePropagation.cpp, src/MultimodePropagation_CPU.cpp,
 * src/RungeKutta.cpp, conventions of src/DiscreteFourierTransform.cpp).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product (libmmprop_hip.so, libmmprop.so, mm_propagation_amd) never links,
 * loads or calls it.
 */
#ifndef MMPROP_ORACLE_H
#define MMPROP_ORACLE_H

#include "../include/mmprop.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_solver orc_solver;

/* Options of the restatement. */
typedef struct orc_options {
  int pair_decode;      /* MMP_PAIRS_REFERENCE_CPU (reference behaviour) or MMP_PAIRS_PHYSICAL */
  int fft_variant;      /* 0: radix-2 decimation in time; 1: radix-2 decimation in frequency */
  int parallel_pairs;   /* 0: serial Raman pair loop as the reference; 1: OpenMP over pairs (same result) */
  int n_threads;        /* OpenMP threads (0: omp_get_max_threads()) */
  int fp32_stages;      /* 0 for the FP64 restatement. Error-budget studies only: ORC_FP32_* bits
                           round the named stage's arithmetic to FP32 (each operation rounded to
                           single precision), to measure what each stage of a mixed-precision
                           solver contributes to its error */
} orc_options;

enum {
  ORC_FP32_FIELDS = 1,  /* U = F^-1[E Fpsi]: the product rounded to FP32, the inverse transform in FP32 */
  ORC_FP32_PAIRS = 2,   /* V_j = U_a conj(U_b) in FP32 */
  ORC_FP32_RAMAN = 4,   /* the Raman convolution (transforms, h_R spectrum product, dt) in FP32 */
  ORC_FP32_GATHER = 8,  /* the Kerr + Raman sum over the nonzeros in FP32 */
  ORC_FP32_SIGMA = 16,  /* F sigma in FP32 (widened to FP64 before the conj(E) product) */
  ORC_FP32_KEEP_CONSTANTS = 32, /* with the bits above: twiddles, coupling values, the h_R spectrum and
                                   the slot-mix factors stay FP64 (only operation results are rounded),
                                   to separate the rounding of the constants from that of the arithmetic */
  ORC_FP32_GATHER_ACC64 = 64  /* with ORC_FP32_GATHER: the products in FP32, their sum over the nonzeros in FP64 */
};

/* Constructor (src/MultimodePropagation.cpp:4-12 + MultimodePropagationCPU_RamanON/OFF
 * constructors, src/MultimodePropagation_CPU.cpp:4-21,108,170-177). NULL on invalid input;
 * orc_last_error() then says why. */
orc_solver* orc_create(const mmp_params* in, const orc_options* opt);
void orc_destroy(orc_solver* s);
const char* orc_last_error(void);

/* computeLawsonRK(nz) (src/MultimodePropagation_CPU.cpp:24-34). */
void orc_compute_lawson_rk(orc_solver* s, unsigned nz);
/* floor(z_final / delta_z) (src/MultimodePropagation.cpp:213-216). */
unsigned orc_default_steps(const orc_solver* s);
/* Steps nz times from the current state without restarting (no reference counterpart;
 * used to check the resumable device entry point). */
void orc_advance(orc_solver* s, unsigned nz);
/* getResult() (src/MultimodePropagation_CPU.cpp:88-104); out is [M][NT] complex. */
void orc_get_result(orc_solver* s, double* out);
/* psi [M][NT] complex and z. */
void orc_get_state(const orc_solver* s, double* psi, double* z);
void orc_set_state(orc_solver* s, const double* psi);

/* Setup products, for pinning against the compiled reference setup. */
typedef struct orc_setup_info {
  unsigned M, nt, stages, n_pairs, n_beta;
  double t_final, delta_t, z_final, delta_z, z, conversion_factor, sign_gamma, T0, ld, max_beta2, max_q;
} orc_setup_info;
void orc_get_setup(const orc_solver* s, orc_setup_info* info);
/* RK tableau: a is [stages][stages] (row i holds a_ij for j < i, zero elsewhere), b, c [stages]. */
void orc_get_tableau(const orc_solver* s, double* a, double* b, double* c);
void orc_get_beta_tilde(const orc_solver* s, double* out);                /* [M][n_beta] */
void orc_get_q_tilde(const orc_solver* s, double* out);                   /* [nnz] in input order */
void orc_get_unique_pairs(const orc_solver* s, unsigned* codes);          /* [n_pairs] */
void orc_get_pair_table(const orc_solver* s, unsigned* a, unsigned* b);   /* [n_pairs] decoded pairs */
void orc_get_nnz_slots(const orc_solver* s, unsigned* slots);             /* [nnz] slot of each nonzero */
/* After orc_init_operators(): L, E1, E[i] exactly as initLawsonRK leaves them. */
void orc_init_operators(orc_solver* s);
void orc_get_linear_operator(const orc_solver* s, double* L);             /* [M][NT] complex */
void orc_get_e1(const orc_solver* s, double* E1);                         /* [M][NT] complex */
void orc_get_e(const orc_solver* s, unsigned stage, double* E);           /* [M][NT] complex */
void orc_get_hr_fft(const orc_solver* s, double* out);                    /* [NT] complex */

/* One nonlinearity evaluation N_i = N(E_i, psi_i) for stage `stage` with the current E
 * tables (src/MultimodePropagation_CPU.cpp:111-166,180-238). psi_i, N_out: [M][NT] complex. */
void orc_nonlinearity(orc_solver* s, unsigned stage, const double* psi_i, double* N_out);

/* Primitives. */
/* DFT with the FFTW conventions used by the reference (src/DiscreteFourierTransform.cpp:12,22-38):
 * sign -1 forward unnormalised, +1 inverse including the 1/n factor. In-place safe. */
void orc_dft(unsigned n, const double* in, double* out, int sign, int variant);

#ifdef __cplusplus
}
#endif

#endif
