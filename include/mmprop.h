/*
 * mmprop.h -- C ABI of the MI355X (gfx950) Lawson-Runge-Kutta multimode
 * propagation library (libmmprop_hip.so).
 *
 * The reference exposes this hot path only as a C++ class API
 * (include/MultimodePropagation.hpp:16-94, include/MultimodePropagation_GPU.cuh:10-85).
 * Every entry point below names the reference member it replaces. Plain
 * pointers and sizes only; no C++ exception crosses this boundary. All
 * arrays are host arrays owned by the caller; the library copies what it
 * needs. Complex arrays are interleaved doubles: element (p, j) of an
 * M x NT complex container lives at [2*(p*NT + j)] (re) and [+1] (im).
 *
 * Threading: one solver handle is used by one host thread at a time.
 * Different handles may be driven from different threads; each context owns
 * its own HIP stream.
 */
#ifndef MMPROP_H
#define MMPROP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MMPROP_ABI_VERSION 6

/* Only the entry points below are exported; the library is built with hidden visibility. */
#if defined(__GNUC__) || defined(__clang__)
#define MMP_API __attribute__((visibility("default")))
#else
#define MMP_API
#endif

typedef enum mmp_status {
  MMP_OK = 0,
  MMP_ERR_INVALID_ARGUMENT = 1, /* malformed parameters (sizes, indices, NULLs, max|Q| == 0, ...) */
  MMP_ERR_UNSUPPORTED = 2,      /* valid for the reference, not supported here (e.g. NT not a power of two) */
  MMP_ERR_OUT_OF_MEMORY = 3,    /* device or host allocation failed */
  MMP_ERR_DEVICE = 4,           /* a HIP call failed */
  MMP_ERR_NO_DEVICE = 5,        /* no usable gfx950 device / kernels not loadable */
  MMP_ERR_NONFINITE = 6,        /* non-finite value in inputs or in the propagated field */
  MMP_ERR_STATE = 7             /* call out of order (e.g. result requested before any step) */
} mmp_status;

/* Precision of the propagation work (FFTs, Kerr/Raman sums). */
enum {
  MMP_PRECISION_FP64 = 0, /* plain FP64 solver: the drop-in for the reference */
  MMP_PRECISION_MIXED = 1 /* FFTs, pair products, Raman convolutions and the Kerr sum in FP32 (twiddles
                             computed in FP64 and rounded); the spectral state, the Lawson phases
                             exp(c_i h L) (the large accumulated beta0 z + beta1 w z phase), the RK stage
                             combination and accumulation, getResult and the energy stay FP64 */
};

/* Which (a, b) pair a Raman/Kerr slot j holds (V_j = U_a * conj(U_b)).
 * REFERENCE_CPU reproduces src/MultimodePropagation_CPU.cpp:126-133,196-204:
 * a = j % M, b = j / M, j being the slot counter. PHYSICAL decodes the stored
 * pair code l + m*M (src/MultimodePropagation.cpp:139-155): a = l, b = m. */
enum {
  MMP_PAIRS_REFERENCE_CPU = 0,
  MMP_PAIRS_PHYSICAL = 1
};

/* Option flags. */
enum {
  MMP_FLAG_NONE = 0,
  MMP_FLAG_PROFILE = 1u << 0,     /* record per-kernel HIP event times (see mmp_kernel_times) */
  MMP_FLAG_NO_GRAPH = 1u << 1,    /* launch step kernels one by one instead of replaying a HIP graph */
  MMP_FLAG_SEPARATE_COLUMNS = 1u << 2, /* keep every column pass a kernel of its own (7 passes per stage)
                                          where the column-stage kernels would run (4 passes per stage:
                                          Raman on, 64-sample columns, n_modes <= 16) */
  MMP_FLAG_AUTOTUNE = 1u << 3 /* at creation, time the column-stage gather variants (wavefronts per
                                 workgroup, Raman lines per LDS chunk) on this solver's own data and
                                 keep the fastest; the choice is cached per structure for the process.
                                 The solver is left at z = 0 with its initial field. Results stay within
                                 rounding of the untuned solver (the sum order over slots changes). */,
  MMP_FLAG_NO_MATRIX_CORES = 1u << 4, /* mixed precision: keep the column-stage gather's coupling sums on
                                         the vector ALUs. By default they run as FP32 matrix products on
                                         the matrix cores at 4 to 10 and 13 to 16 modes (the same
                                         arithmetic, another order of FP32 rounding), and the tuner tries
                                         both. */
  MMP_FLAG_MATRIX_CORES = 1u << 5     /* mixed precision: the coupling sums on the matrix cores wherever
                                         that gather exists (4 to 16 modes with parity-compact coupling
                                         blocks, 4 to 10 without), also where the default keeps them on
                                         the vector ALUs (11 and 12 modes); the tuner then tries only such
                                         variants. Not with MMP_FLAG_NO_MATRIX_CORES. The default does not
                                         depend on the number of fibers, so a batch and its fibers solved
                                         alone run the same arithmetic. */
};

/*
 * The reference's SimulationParameters (include/MultimodePropagation.hpp:16-33,
 * called InputParameters in README.md:71-86), flattened.
 *
 * coupling_coefficients: Q^(p) for p < n_modes is the COO list
 *   entries [coupling_offsets[p], coupling_offsets[p+1]) of coupling_klm (k,l,m
 *   triplets) and coupling_values, in insertion order
 *   (src/Sparse3DArray.cpp:16-19). The order matters: it fixes the pair-slot
 *   order (src/MultimodePropagation.cpp:139-155).
 * raman: 1 selects the RamanON class (raman_response required), 0 the RamanOFF
 *   class (raman_proportion and raman_response ignored), mirroring the class
 *   choice of matlab/mm_propagation.cpp:59.
 */
typedef struct mmp_params {
  uint32_t n_modes;                       /* M */
  double fiber_length;                    /* m */
  uint32_t n_beta;                        /* dispersion orders per mode, >= 3 */
  const double* dispersion_coefficients;  /* [M][n_beta], ps^l/m */
  const uint32_t* coupling_offsets;       /* [M+1] */
  const uint32_t* coupling_klm;           /* [nnz][3] */
  const double* coupling_values;          /* [nnz], m^-2 */
  uint32_t nt;                            /* NT */
  double time_window;                     /* ps */
  double pulse_width;                     /* T0, ps */
  const double* initial_fields;           /* [M][NT] complex, W^(1/2) */
  uint32_t n_steps;
  uint32_t method_order;                  /* 8 -> 11 stages, 6 -> 7 stages, else RK4 */
  double nonlinearity_const;              /* gamma, W^-1 m */
  double raman_proportion;                /* fR */
  const double* raman_response;           /* [NT], h_R sampled on t_j = j*dt */
  uint32_t raman;                         /* 0 = RamanOFF, 1 = RamanON */
} mmp_params;

typedef struct mmp_options {
  uint32_t precision;    /* MMP_PRECISION_* */
  uint32_t pair_decode;  /* MMP_PAIRS_* */
  uint32_t flags;        /* MMP_FLAG_* */
  uint32_t reserved;
} mmp_options;

/* Scalars the reference computes in its constructor
 * (src/MultimodePropagation.cpp:31-63, 139-155) plus the solver state. */
typedef struct mmp_info {
  uint32_t n_modes, nt, stages, n_pairs, nnz, n_conv_lines, n1, n2;
  uint32_t steps_taken;      /* steps since the last restart (z = 0) */
  uint32_t default_steps;    /* floor(z_final / delta_z), src/MultimodePropagation.cpp:213-216 */
  double t_final, delta_t, z_final, delta_z, z;
  double conversion_factor, sign_gamma, ld, max_beta2, max_q;
  size_t device_bytes;       /* device memory held by the solver (all fibers) */
  uint32_t n_fibers;         /* fibers in the solver (1 unless made by mmp_batch_create) */
  uint32_t live;             /* 1 when the steps advance this fiber, 0 when it is parked */
  uint32_t gather_waves;     /* column-stage gather: wavefronts per workgroup (0: separate passes) */
  uint32_t chunk_lines;      /* column-stage gather: Raman lines per LDS chunk at most */
  uint32_t tuned;            /* 1 when MMP_FLAG_AUTOTUNE chose the two above */
  uint32_t pair_columns;     /* column-stage pair pass: neighbouring columns per workgroup (0: separate passes) */
  uint32_t pair_groups;      /* column-stage pair pass: workgroups sharing one column block's Raman lines */
  uint32_t gather_hi_occupancy; /* 1 when the gather runs its three-workgroups-per-CU variant */
  uint32_t gather_matrix_cores; /* 1 when the gather's coupling sums run on the matrix cores (FP32) */
  uint32_t gather_chunks;    /* column-stage gather: LDS chunks of Raman lines and slots per launch */
  uint32_t halted;           /* 1 when the device parked this fiber because its field stopped being finite
                                (after the step that made it so: z and steps_taken count that step) */
} mmp_info;

typedef struct mmp_ctx mmp_ctx;
typedef struct mmp_solver mmp_solver;

/* Status text for a status code; never NULL. */
MMP_API const char* mmp_status_string(mmp_status status);

/* Detail of the last failed call on this host thread ("" if none). */
MMP_API const char* mmp_last_error(void);

/* ABI version compiled into the library (MMPROP_ABI_VERSION). */
MMP_API int mmp_abi_version(void);

/* Digest of the sources the library was built from (the first 16 hex digits of a SHA-256 over the
 * library's source files and this header, as mm_propagation_amd/build.py computes it), so that a
 * caller can tell a library built from other sources; "unknown" when the build did not set it. */
MMP_API const char* mmp_source_hash(void);

/* Host-only planning (no device needed, no reference counterpart): validates
 * params and options exactly as mmp_solver_create does and fills the solver
 * scalars and the device memory a solver would hold (device_bytes). Used for
 * admission by memory footprint. */
MMP_API mmp_status mmp_plan(const mmp_params* params, const mmp_options* options, mmp_info* info_out);

/* Host-only: the prepared arrays of mmp_plan -- the dimensionless initial field
 * psi0 = conversion_factor * A0 ([M][NT] complex, src/MultimodePropagation.cpp:59-63),
 * Im L_p(alpha_j) in natural order ([M][NT], src/MultimodePropagation.cpp:165-187),
 * the decoded pair (a_j, b_j) of every slot ([n_pairs] each), and the initial Lawson
 * tables E1 = exp(h L) ([M][NT] complex) and E_i = exp(c_i h L) ([stages][M][NT] complex)
 * of src/MultimodePropagation.cpp:190-199, in natural order. Any pointer may be NULL. */
MMP_API mmp_status mmp_plan_arrays(const mmp_params* params, const mmp_options* options, double* psi0_out,
                           double* imL_out, uint32_t* pair_a_out, uint32_t* pair_b_out, double* e1_out,
                           double* e_out);

/* Host-only: the Butcher tableau the solver uses for method_order
 * (src/MultimodePropagation.cpp:15-27 with src/RungeKutta.cpp:3-72,125-175):
 * *stages_out = s (4, 7 or 11); a_out is [11][11] row-major (a_ij for j < i, zero
 * elsewhere), b_out and c_out are [11]. */
MMP_API mmp_status mmp_tableau(uint32_t method_order, uint32_t* stages_out, double* a_out, double* b_out,
                               double* c_out);

/* Device context: one HIP stream on one device. No reference counterpart:
 * the reference hard-codes device 0 and the default stream
 * (src/MultimodePropagation_GPU.cu:17). */
MMP_API mmp_status mmp_ctx_create(int device, mmp_ctx** out);
MMP_API void mmp_ctx_destroy(mmp_ctx* ctx);

/* Replaces the constructors MultimodePropagationGPU_RamanON/OFF(const
 * SimulationParameters&) (include/MultimodePropagation_GPU.cuh:184,196;
 * src/MultimodePropagation.cpp:4-12): validates and copies the inputs, makes
 * them dimensionless, builds the pair table and uploads everything once.
 * options may be NULL (FP64, REFERENCE_CPU pairs, graphs on). */
MMP_API mmp_status mmp_solver_create(mmp_ctx* ctx, const mmp_params* params,
                             const mmp_options* options, mmp_solver** out);
MMP_API void mmp_solver_destroy(mmp_solver* solver);

/* Batched solver (no reference counterpart: the reference runs one fiber per object).
 * n_fibers independent fibers, params[f] each, that share n_modes, nt, method_order, raman and
 * the pair-slot structure of their coupling lists (their values, dispersion, lengths, pulses
 * and Raman responses may all differ). Every stage pass runs all fibers in one launch. Needs
 * the column-stage kernels (Raman on, nt in [4096, 65536], n_modes <= 16). The calls
 * above act on every fiber (mmp_compute_lawson_rk, mmp_advance, mmp_propagate_snapshots's
 * stepping) or on fiber 0 (results, state, energy, info); the mmp_fiber_* calls address one
 * fiber. */
MMP_API mmp_status mmp_batch_create(mmp_ctx* ctx, uint32_t n_fibers, const mmp_params* params,
                                    const mmp_options* options, mmp_solver** out);
MMP_API mmp_status mmp_fiber_result(mmp_solver* solver, uint32_t fiber, double* fields_out);
MMP_API mmp_status mmp_fiber_state(mmp_solver* solver, uint32_t fiber, double* psi_out, double* z_out);
/* Sets fiber f's dimensionless field and restarts that fiber alone at z = 0. */
MMP_API mmp_status mmp_fiber_set_state(mmp_solver* solver, uint32_t fiber, const double* psi);
MMP_API mmp_status mmp_fiber_energy(mmp_solver* solver, uint32_t fiber, double* energy_out);
MMP_API mmp_status mmp_fiber_info(mmp_solver* solver, uint32_t fiber, mmp_info* info_out);
/* Replaces fiber f by a new problem of the same structure (a serving slot taking a new
 * request): its inputs are uploaded, it restarts at z = 0; the other fibers keep their state.
 * Waits for the work already queued on the context's stream. */
MMP_API mmp_status mmp_fiber_load(mmp_solver* solver, uint32_t fiber, const mmp_params* params);
/* Parks (live = 0) or resumes (live = 1) fiber f: the steps leave a parked fiber's field, z and
 * Lawson tables untouched, and its workgroups exit at once, so a batch whose slots are partly
 * empty costs about what its live fibers cost. mmp_fiber_load makes a fiber live. Waits for the
 * work already queued on the context's stream. Fibers start live. MMP_ERR_UNSUPPORTED for a
 * solver that does not run the column-stage kernels (batches always do).
 * A fiber whose spectral state stops being finite is parked by the device itself at the end of
 * that step (no reference counterpart: the reference keeps stepping a NaN field): its later steps
 * cost nothing, its z and step count stop there, mmp_info.halted reports it, and resuming it with
 * live = 1 returns MMP_ERR_NONFINITE; mmp_fiber_load or mmp_fiber_set_state gives it a finite
 * field and makes it live again. */
MMP_API mmp_status mmp_fiber_set_live(mmp_solver* solver, uint32_t fiber, uint32_t live);

/* Device memory of one mmp_solve_static call: every device array of the solve must fit in it. */
#define MMP_STATIC_SCRATCH_BYTES ((size_t)64 << 20)

/* No-allocation solve. No counterpart in the reference: for callers that cannot allocate device
 * memory (code running under HIP graph capture, tight serving loops, hosts that forbid hipMalloc).
 * Equivalent to mmp_solver_create(ctx, params, options) + mmp_compute_lawson_rk(nz) +
 * mmp_get_result(fields_out) + mmp_solver_destroy, with the same kernels and the same results bit
 * for bit, but every device array lives in a static per-device scratch of MMP_STATIC_SCRATCH_BYTES
 * that the library holds until the process ends: the call makes no hipMalloc, hipFree, graph
 * instantiation or event. Kernels are launched one by one on the context's stream; the call returns
 * once the result is on the host. The scratch keeps the last call's contents (each call uploads
 * everything it reads). Calls are serialised by a process-wide lock: calls from several threads
 * wait for each other, since overlapping calls on one device would race on the one scratch.
 * MMP_ERR_UNSUPPORTED when the problem's device footprint (mmp_plan's device_bytes plus alignment)
 * exceeds the scratch; MMP_ERR_INVALID_ARGUMENT with MMP_FLAG_PROFILE or MMP_FLAG_AUTOTUNE. */
MMP_API mmp_status mmp_solve_static(mmp_ctx* ctx, const mmp_params* params, const mmp_options* options, uint32_t nz,
                                    double* fields_out);

/* Solvers of different shapes advanced together (no reference counterpart: the reference runs one
 * fiber per object). A group holds solvers and batches of any n_modes, nt and precision, made on
 * the device of ctx with HIP graphs (neither MMP_FLAG_PROFILE nor MMP_FLAG_NO_GRAPH); one graph,
 * captured at mmp_group_create, holds 16 steps of every member, so mmp_group_advance(g, n) steps
 * all of them with one graph launch per 16 steps (and one per remaining step) on ctx's stream.
 * Members made on ctx run in order; members made on contexts of their own are parallel branches of
 * the graph (forked from ctx's stream and joined back), so their kernels may overlap on the device.
 * It is mmp_advance(member, n) for every member, bit for bit, and each member's stream is ordered
 * after the group's work. The members keep their own calls; destroy the group before them. */
typedef struct mmp_group mmp_group;
MMP_API mmp_status mmp_group_create(mmp_ctx* ctx, uint32_t n_solvers, mmp_solver* const* solvers, mmp_group** out);
MMP_API mmp_status mmp_group_advance(mmp_group* group, uint32_t n_steps);
MMP_API void mmp_group_destroy(mmp_group* group);

/* Replaces computeLawsonRK(nz) (include/MultimodePropagation_GPU.cuh:143;
 * semantics of src/MultimodePropagation_CPU.cpp:24-34): restarts at z = 0
 * with freshly built Lawson operators, keeps the current field, then takes
 * nz steps. */
MMP_API mmp_status mmp_compute_lawson_rk(mmp_solver* solver, uint32_t nz);

/* Replaces computeLawsonRK() (src/MultimodePropagation.cpp:213-216):
 * mmp_compute_lawson_rk with nz = floor(z_final / delta_z). */
MMP_API mmp_status mmp_compute_lawson_rk_default(mmp_solver* solver);

/* Resumable stepping (no reference counterpart): continues from the current
 * z without restarting. mmp_compute_lawson_rk(s, a) followed by
 * mmp_advance(s, b) equals mmp_compute_lawson_rk(s, a + b). */
MMP_API mmp_status mmp_advance(mmp_solver* solver, uint32_t n_steps);

/* Snapshot modes of mmp_propagate_snapshots. */
enum {
  MMP_SNAPSHOTS_REFERENCE_RESTART = 0, /* each chunk is computeLawsonRK(nz): z and the Lawson tables
                                          restart per chunk, as matlab/mm_propagation.cpp:61-101 does */
  MMP_SNAPSHOTS_RESUMABLE = 1          /* one solve from z = 0, paused at each snapshot (mmp_advance) */
};

/* Replaces the n_save loop of the MEX gateway (matlab/mm_propagation.cpp:61-101): n_steps split
 * into n_save chunks of n_steps / n_save steps, the last one taking the remainder; after each
 * chunk the physical field (mmp_get_result) is written to fields_out[i] ([n_save][M][NT] complex)
 * and the distance Z to z_out[i] in metres, accumulated as Z += nz * (fiber_length / n_steps)
 * like the gateway. Either output may be NULL. n_save must be in [1, n_steps]. */
MMP_API mmp_status mmp_propagate_snapshots(mmp_solver* solver, uint32_t n_steps, uint32_t n_save, uint32_t mode,
                                           double* fields_out, double* z_out);

/* Replaces getResult() (src/MultimodePropagation_CPU.cpp:88-104): the
 * physical field A_p = F^-1[exp(L_p z) F psi_p] / conversion_factor, written
 * to fields_out [M][NT] complex. */
MMP_API mmp_status mmp_get_result(mmp_solver* solver, double* fields_out);

/* Dimensionless interaction-picture field psi [M][NT] complex and the
 * accumulated z (the reference's protected psi and z members). Either
 * pointer may be NULL. */
MMP_API mmp_status mmp_get_state(mmp_solver* solver, double* psi_out, double* z_out);

/* Replaces the dimensionless field and restarts at z = 0 (the reference
 * keeps psi across computeLawsonRK calls; this sets it explicitly).
 * psi is [M][NT] complex, in dimensionless units. */
MMP_API mmp_status mmp_set_state(mmp_solver* solver, const double* psi);

/* Per-mode energy of the physical field, E_p = sum_j |A_p,j|^2 * (T_w / NT)
 * in W*ps, computed on the device in FP64. energy_out is [M]. */
MMP_API mmp_status mmp_energy(mmp_solver* solver, double* energy_out);

/* Per-mode photon number of the physical field (no reference counterpart; the reference writes
 * only fields, matlab/mm_propagation.cpp:72-75, simulations/main.cpp:101-103):
 * N_p = sum_k E_p,k / (hbar * omega_k), with E_p,k = (T_w / NT^2) |(F A_p)_k|^2 the energy of
 * frequency bin k (W ps; F the unnormalised forward DFT) and omega_k = omega0 - 2 pi k / T_w its
 * absolute angular frequency, k in (-NT/2, NT/2] as the reference's frequency grid
 * (src/MultimodePropagation.cpp:169-175) and omega0 = 2 pi c / wavelength. The carrier wavelength
 * (metres) is an argument because SimulationParameters has none. Computed on the device in FP64
 * from the spectral state. photons_out is [M]. MMP_ERR_INVALID_ARGUMENT when wavelength is not
 * positive or the grid reaches omega_k <= 0 (omega0 <= pi NT / T_w). */
MMP_API mmp_status mmp_photon_number(mmp_solver* solver, double wavelength, double* photons_out);
MMP_API mmp_status mmp_fiber_photon_number(mmp_solver* solver, uint32_t fiber, double wavelength,
                                           double* photons_out);

/* Solver scalars and state. */
MMP_API mmp_status mmp_solver_info(mmp_solver* solver, mmp_info* info_out);

/* With MMP_FLAG_PROFILE: the names and summed HIP-event durations (ms) and
 * launch counts of the kernels issued since the last call (then reset).
 * Writes up to capacity entries; *count_out receives the number available. */
typedef struct mmp_kernel_time {
  char name[48];
  double total_ms;
  uint64_t launches;
  double bytes_per_launch;  /* algorithmic HBM bytes of one launch */
} mmp_kernel_time;
MMP_API mmp_status mmp_kernel_times(mmp_solver* solver, mmp_kernel_time* out,
                            uint32_t capacity, uint32_t* count_out);

/* Block until all work queued on the context's stream is complete. */
MMP_API mmp_status mmp_ctx_synchronize(mmp_ctx* ctx);

/* Without blocking: MMP_OK when all work queued on the context's stream is complete,
 * MMP_ERR_STATE while some is still running (no reference counterpart; for host loops that
 * keep several contexts busy). */
MMP_API mmp_status mmp_ctx_query(mmp_ctx* ctx);

/* ------------------------------------------------------------------------------------------
 * Request server (no reference counterpart): continuous batching of propagation solves of
 * mixed n_modes, nt and step counts on one device. A request is admitted when its device
 * footprint (mmp_plan) fits the budget: into a free slot of the batch group of its structure
 * (requests the batched kernels take; mmp_fiber_load at the group's next chunk boundary) or a
 * solver of its own. Every group and solo solver has its own stream, so their work overlaps;
 * a group advances by the fewest steps any of its live requests still needs (at most
 * max_chunk), finished requests are read back and their slots parked. Each result is
 * bit-identical to the same solve alone with the same options, unless the flags hold
 * MMP_FLAG_AUTOTUNE: a group then runs the gather variant tuned for its slot count, which may
 * sum over the slots in another order than the solve alone, and results agree to rounding
 * (about 1e-15 relative in FP64). One host thread drives a server; mmp_serve_poll never blocks.
 * ---------------------------------------------------------------------------------------- */
typedef struct mmp_server mmp_server;

typedef struct mmp_serve_options {
  int device;
  uint32_t slots;        /* fiber slots per batch group, >= 1 */
  uint32_t max_chunk;    /* steps a group or solver takes between scheduling points, >= 1 */
  uint32_t precision;    /* MMP_PRECISION_* for every solver */
  uint32_t pair_decode;  /* MMP_PAIRS_* */
  uint32_t flags;        /* MMP_FLAG_* for every solver (MMP_FLAG_AUTOTUNE tunes each structure once;
                            results are then rounding-equal, not bit-identical, to solves alone) */
  double hbm_budget;     /* bytes of device memory the server's solvers may hold */
} mmp_serve_options;

enum { MMP_REQ_QUEUED = 0, MMP_REQ_RUNNING = 1, MMP_REQ_DONE = 2, MMP_REQ_NONFINITE = 3 };

typedef struct mmp_serve_record {
  uint32_t state;      /* MMP_REQ_* */
  uint32_t batched;    /* 1: a slot of a batch group, 0: a solver of its own */
  uint32_t group;      /* the group's index when batched */
  uint32_t slot;       /* the slot when batched */
  double t_submit;     /* seconds since mmp_serve_create */
  double t_admit;
  double t_finish;
  uint64_t units;      /* n_modes * nt * n_steps */
  size_t device_bytes; /* planned footprint of one fiber of this request */
} mmp_serve_record;

typedef struct mmp_serve_stats {
  uint64_t submitted, queued, running, done, nonfinite;
  uint32_t groups, solos;       /* live now */
  uint32_t groups_created, solos_created;
  size_t bytes_in_use, peak_bytes;
} mmp_serve_stats;

MMP_API mmp_status mmp_serve_create(const mmp_serve_options* options, mmp_server** out);
MMP_API void mmp_serve_destroy(mmp_server* server);
/* Copies the request (every array of params); it is queued at the current time. The request's
 * n_steps argument is the number of steps to take (params->n_steps sets h = L / n_steps). */
MMP_API mmp_status mmp_serve_submit(mmp_server* server, const mmp_params* params, uint32_t n_steps,
                                    uint64_t* id_out);
/* One scheduling pass: harvests groups and solvers whose chunk has finished, admits queued
 * requests that fit, launches the next chunks. active_out (optional) = queued + running. */
MMP_API mmp_status mmp_serve_poll(mmp_server* server, uint32_t* active_out);
/* Polls until no request is queued or running. MMP_ERR_STATE if a queued request can never be
 * admitted (its footprint exceeds the budget) while nothing runs. */
MMP_API mmp_status mmp_serve_drain(mmp_server* server);
MMP_API mmp_status mmp_serve_record_get(mmp_server* server, uint64_t id, mmp_serve_record* out);
/* The physical field of a finished request ([M][NT] complex); MMP_ERR_STATE while it is queued
 * or running, MMP_ERR_NONFINITE when its field is not finite. */
MMP_API mmp_status mmp_serve_result(mmp_server* server, uint64_t id, double* fields_out);
MMP_API mmp_status mmp_serve_stats_get(mmp_server* server, mmp_serve_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* MMPROP_H */
