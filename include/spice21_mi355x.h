/*
 * spice21_mi355x.h -- C ABI of the MI355X-native Spice21 Newton engine.
 *
 * The engine replaces the bodies of Spice21's Newton inner loop: component loading and matrix
 * stamping (spice21/src/analysis.rs:153-168), the residual (spice21/src/sparse21/mod.rs:298-327),
 * and LU factorisation plus forward/backward solves (spice21/src/sparse21/mod.rs:647-991), for
 * batches of circuit instances resident on one MI355X.
 *
 * Every entry point is synchronous, never unwinds across the ABI, returns an S21_* status
 * and leaves the reference's error text (SpError.desc, spice21/src/spresult.rs:9-48) in
 * s21_last_error().  No torch or HIP types appear in any signature.
 */
#ifndef SPICE21_MI355X_H
#define SPICE21_MI355X_H

#ifdef __HIPCC_RTC__
#include <cstdint> /* runtime-compiled kernels see only the type and status declarations */
#else
#include <stddef.h>
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define S21_ABI_VERSION 2
#define S21_GND (-1)
#define S21_API __attribute__((visibility("default")))

/* ---- status codes ------------------------------------------------------------------------ */
enum {
    S21_OK = 0,
    S21_ERR = 1,          /* any other SpError; text in s21_last_error()                        */
    S21_CONVERGENCE = 2,  /* "Convergence Failed"        analysis.rs:209 (real), :302 (complex)  */
    S21_SINGULAR = 3,     /* "Singular Matrix"           sparse21/mod.rs:651,656,868,957,971     */
    S21_PIVOT = 4,        /* "Pivot Search Fail" / zero pivot  sparse21/mod.rs:665,872          */
    S21_DECODE = 5,       /* "Decode Error: ..."         proto.rs:15-19                          */
    S21_INVALID = 6,      /* malformed argument (null pointer, bad index, size beyond limits)    */
    S21_DEVICE = 7,       /* HIP runtime failure or no gfx950 device                             */
    S21_NONFINITE = 8     /* reserved, never reported: a NaN passes the convergence test here as in
                             the reference (analysis.rs:331-345 compares with >, false for NaN)    */
};

/* ---- flattened netlist (the output of elaboration, elab.rs:244-265) ----------------------- */
enum { S21_VAR_V = 0, S21_VAR_I = 1 };                       /* analysis.rs:33-38 */

enum {                                                      /* comps/mod.rs:60-72 */
    S21_COMP_R = 0,      /* Resistor, value = conductance g          comps/mod.rs:271-323 */
    S21_COMP_C = 1,      /* Capacitor, value = c                      comps/mod.rs:152-239 */
    S21_COMP_I = 2,      /* Isrc, value = dc                          comps/mod.rs:326-346 */
    S21_COMP_V = 3,      /* Vsrc, value = dc, acm = AC magnitude      comps/mod.rs:95-150  */
    S21_COMP_MOS1 = 4,   /* Mos level 1                               comps/mos.rs:625-969 */
    S21_COMP_MOS0 = 5,   /* Mos level 0 (model = 0 NMOS / 1 PMOS)     comps/mos.rs:1007-1100 */
    S21_COMP_DIODE = 6,  /* Diode                                     comps/diode.rs:215-355 */
    S21_COMP_BSIM4 = 7   /* BSIM4 (comps/bsim4/bsim4solver.rs:3757-3769): operating points and
                            transients, in batches and on the single-instance path; no AC
                            (the reference's BSIM4 has no load_ac)                        */
};

/* Optional-double parameter set keyed by proto field number (protos/mos.proto:22-76,
 * protos/spice21.proto:43-68): v[f] is used iff bit f of `set` is 1, otherwise the reference
 * default applies.  Enum and integer fields (Mos1Model.mos_type = 2, tpg = 28) are stored as
 * doubles. */
#define S21_PARAM_SLOTS 40
typedef struct {
    double v[S21_PARAM_SLOTS];
    uint64_t set;
} s21_params;

typedef struct {
    int32_t type;     /* S21_COMP_* */
    int32_t node[6];  /* variable indices, S21_GND for ground:
                         R, C, I: {p, n}; V: {p, n, ivar}; MOS1: {d, g, s, b, dp, sp};
                         MOS0: {d, g, s, b}; DIODE: {p, n, r}; BSIM4: {d, g, s, b}     */
    int32_t model;    /* MOS1: index into mos1_models; DIODE: into diode_models; MOS0: 0|1;
                         BSIM4: into bsim4_models                                          */
    int32_t inst;     /* MOS1: index into mos1_insts; DIODE: into diode_insts;
                         BSIM4: into bsim4_insts                                           */
    double value;     /* R: g; C: c; I: dc; V: dc                                         */
    double acm;       /* V: AC magnitude                                                  */
} s21_comp;

typedef struct {
    int32_t n_vars;
    const int32_t* var_kind;          /* [n_vars] S21_VAR_*                              */
    const char* const* var_name;      /* [n_vars] signal names, may be NULL              */
    int32_t n_comps;
    const s21_comp* comps;            /* [n_comps] in elaboration order                  */
    int32_t n_mos1_models;  const s21_params* mos1_models;
    int32_t n_mos1_insts;   const s21_params* mos1_insts;
    int32_t n_diode_models; const s21_params* diode_models;
    int32_t n_diode_insts;  const s21_params* diode_insts;
    /* Bsim4Model (protos/bsim4.proto:44-46: slot 1 = mos_type, 1 PMOS) and Bsim4InstParams keyed by
     * field number (protos/bsim4.proto:6-41); s21_override field numbers on a BSIM4 start at 1.  Only the polarity reaches the model
     * (Bsim4ModelSpecs::from, comps/bsim4/model/mod.rs:21-30); every mode selector is at its default,
     * so a BSIM4 device has no internal nodes (bsim4ports.rs:31-112). */
    int32_t n_bsim4_models; const s21_params* bsim4_models;
    int32_t n_bsim4_insts;  const s21_params* bsim4_insts;
} s21_netlist;

/* Simulation options, defaults as analysis.rs:665-693:
 * temp = tnom = 300.15 K, gmin = 1e-12 S, iabstol = 1e-12 A, reltol = 1e-3. */
typedef struct {
    double temp, tnom, gmin, iabstol, reltol;
} s21_options;

/* Transient options (analysis.rs:451-456).  Initial conditions force node ic_var[k] to ic_val[k]
 * through a 1 S resistor and a Vsrc during the operating point, then 1e-9 S / 0 V for the rest
 * of the run (analysis.rs:510-549); ic_name[k] names the appended ".{name}.vic" / ".{name}.iic"
 * variables. */
typedef struct {
    double tstep, tstop;
    int32_t n_ic;
    const int32_t* ic_var;
    const double* ic_val;
    const char* const* ic_name;
} s21_tran_options;

/* AC options (analysis.rs:700-714): npts is the total count, not per decade. */
typedef struct {
    uint64_t fstart, fstop, npts;
} s21_ac_options;

/* ---- library / errors --------------------------------------------------------------------- */
S21_API int32_t s21_abi_version(void);
/* Copies the text of the most recent error of the calling thread; returns its full length. */
S21_API size_t s21_last_error(char* buf, size_t len);
S21_API void s21_default_options(s21_options* opts);
/* Number of time points the reference's fixed-step loop produces, including the OP point at
 * t = 0 (analysis.rs:551-570: t = tstep; while t < tstop { ...; t += tstep }). */
S21_API int64_t s21_tran_point_count(double tstep, double tstop);
/* Number of AC frequencies (analysis.rs:791-819, including both ends). */
S21_API int64_t s21_ac_point_count(const s21_ac_options* ac);

/* ---- circuits and elaboration (circuit.rs, elab.rs) ---------------------------------------- */
typedef struct s21_ckt s21_ckt;
typedef struct s21_elab s21_elab;

/* Decode a protobuf `spice21.Circuit` (protos/spice21.proto:128-133) as Ckt::decode /
 * Ckt::from_proto do (circuit.rs:248-256, 281-330). */
S21_API int s21_ckt_decode(const uint8_t* buf, size_t len, s21_ckt** out);
/* Register a level-0 MOS model name, as the reference's test helper does by inserting into
 * `defs.mos0` (tests.rs:1443-1447); proto carries no Mos0 definition. */
S21_API int s21_ckt_add_mos0(s21_ckt* ckt, const char* name, int32_t mos_type);
S21_API void s21_ckt_free(s21_ckt* ckt);
/* Flatten the hierarchy into variables and components (elab.rs:244-265). */
S21_API int s21_elaborate(const s21_ckt* ckt, s21_elab** out);
/* View of the flattened netlist; valid until s21_elab_free. */
S21_API int s21_elab_netlist(const s21_elab* e, s21_netlist* out);
S21_API void s21_elab_free(s21_elab* e);

/* ---- batched engine -------------------------------------------------------------------------
 * A batch holds one or more groups.  A group is one topology (a flattened netlist) times n_inst
 * instances that differ only in parameter values.  All instances of all groups are solved by one
 * persistent kernel launch per analysis call. */
typedef struct s21_batch s21_batch;

enum { S21_OVR_VALUE = 0, S21_OVR_MODEL = 1, S21_OVR_INST = 2 };
typedef struct {
    int32_t comp;         /* component index in the group's netlist                       */
    int32_t target;       /* S21_OVR_VALUE: comp.value; S21_OVR_MODEL / S21_OVR_INST: a
                             proto field of that component's model / instance params      */
    int32_t field;        /* proto field number for MODEL / INST targets                  */
    const double* vals;   /* [n_inst] per-instance values                                 */
} s21_override;

typedef struct {
    const s21_netlist* net;
    int32_t n_inst;
    int32_t n_over;
    const s21_override* over;
} s21_group;

enum { S21_FP64 = 0, S21_MIXED = 1 };

/* Build per-instance derived parameters (comps/mos.rs:320-476, comps/diode.rs:146-212) and keep
 * the groups resident on the current HIP device.  precision: S21_FP64 factors and solves in
 * double; S21_MIXED factors and solves in float while device evaluation, the KCL residual, the
 * Newton update and the convergence test stay in double. */
S21_API int s21_batch_create(int32_t n_groups, const s21_group* groups, const s21_options* opts,
                     int32_t precision, s21_batch** out);
S21_API void s21_batch_free(s21_batch* b);

/* DC operating point of every instance (analysis.rs:383-388 and Solver::solve :169-210). */
S21_API int s21_batch_op(s21_batch* b);
/* Operating point of every instance as s21_batch_op computes it, allocating nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).
 *   - The batch must already hold its operating-point structure: one earlier s21_batch_op call on it
 *     builds and uploads that (and allocates); otherwise S21_INVALID.
 *   - x [sum over groups of n_inst * n_vars], status and iters [sum over groups of n_inst] are caller-
 *     allocated DEVICE buffers on the batch's device, groups in order and instances in order within a
 *     group.  x's contents are ignored: the solve starts from x = 0 as the reference's does.
 *   - The Newton state lives in a fixed 64 MiB device array of the library (one per GPU, kept until
 *     the process ends, cleared by each call); a batch whose state needs more returns S21_INVALID.
 *     BSIM4 devices keep their points in the batch's own buffers (allocated by the preparing
 *     s21_batch_op), which each call clears.
 *   - An instance whose frozen pivot order meets a pivot it cannot keep (zero, or below the smallest
 *     normal number of the factors' precision) reports S21_PIVOT (s21_batch_op resumes such instances
 *     on a new order in a launch of its own, which allocates).
 *   - The batch's own results (s21_batch_get_*) are left unchanged; s21_batch_timing reports this call.
 *   - Every call on one GPU shares that GPU's array: calls that overlap on one GPU race. */
S21_API int s21_batch_op_noalloc(s21_batch* b, double* x, int32_t* status, int64_t* iters);
/* Fixed-step Backward-Euler transient with initial-condition forcing (analysis.rs:489-574). */
S21_API int s21_batch_tran(s21_batch* b, const s21_tran_options* t);
/* Operating point, then complex small-signal sweep (analysis.rs:761-832). */
S21_API int s21_batch_ac(s21_batch* b, const s21_ac_options* a);

/* The reference's AC result files for one instance of the last sweep, written into `dir` (analysis.rs:
 * 783-827): `stream.ac.json` (per frequency a flat array [f, re0, im0, re1, im1, ...]) and `ac.json` (the
 * serialised AcResult: signals, freq, data as [re, im] pairs, and the map, empty as the reference writes
 * it).  s21_ac_bytes writes them into the working directory after its sweep, as the reference's `ac` does. */
S21_API int s21_batch_write_ac_results(const s21_batch* b, int32_t group, int32_t inst, const char* dir);
S21_API int32_t s21_batch_n_groups(const s21_batch* b);
S21_API int32_t s21_batch_n_inst(const s21_batch* b, int32_t group);
/* Unknowns of the group after the last analysis (includes IC variables after a transient). */
S21_API int32_t s21_batch_n_vars(const s21_batch* b, int32_t group);
/* Name of unknown k of the group after the last analysis (stable until the next analysis). */
S21_API const char* s21_batch_var_name(const s21_batch* b, int32_t group, int32_t k);
/* Final solution of each instance: out[i * n_vars + k]. */
S21_API int s21_batch_get_x(const s21_batch* b, int32_t group, double* out);
/* Per-instance S21_* status and Newton iteration totals of the last analysis. */
S21_API int s21_batch_get_status(const s21_batch* b, int32_t group, int32_t* status, int64_t* iters);
/* Number of stored points of the last transient / AC sweep. */
S21_API int64_t s21_batch_n_points(const s21_batch* b);
/* Time (or frequency) axis of the last transient (AC) analysis: out[npts]. */
S21_API int s21_batch_get_axis(const s21_batch* b, double* out);
/* Waveform of one instance: out[p * n_vars + k] (transient); AC writes interleaved re/im,
 * out[(p * n_vars + k) * 2 + {0,1}]. */
S21_API int s21_batch_get_wave(const s21_batch* b, int32_t group, int32_t inst, double* out);
/* Device time of the last analysis's Newton kernel launches (HIP events, milliseconds), the
 * number of those launches, and the average duration of the dominant kernel. */
S21_API int s21_batch_timing(const s21_batch* b, double* total_ms, int64_t* launches, double* kernel_avg_ms);
/* Resumptions of the last operating point or transient: the instance-launches that continued an instance
 * whose kept pivot order met a pivot it cannot keep (zero, sparse21/mod.rs:872, or below the smallest
 * normal number of the factors' precision), each on the order the reference's search picks there
 * (sparse21/mod.rs:647-919).  0 when every instance kept its order throughout. */
S21_API int s21_batch_resumptions(const s21_batch* b, int64_t* instance_rounds);

/* ---- introspection used by the parity tests ------------------------------------------------ */
/* Host symbolic phase only (no GPU): pivot order (row/col i2e, each [n_vars + 2 * n_ic]) that the
 * engine freezes for this netlist -- the reference's first factorisation of the operating-point
 * Jacobian at x = 0 (sparse21/mod.rs:647-919) -- and the sizes of the resulting structure.  `tran`
 * may be NULL (operating point) or carry the initial conditions of a transient. */
S21_API int s21_symbolic(const s21_netlist* net, const s21_options* opts, const s21_tran_options* tran,
                         int32_t* row_i2e, int32_t* col_i2e, int64_t* nnz_a, int64_t* nnz_lu, int64_t* levels,
                         int64_t* flops_lu);
/* Host only (no GPU): symbolic phase, then the generated factorisation of the frozen pattern compiled
 * for gfx950 without loading it; reports its statements and code-object size. */
S21_API int s21_jit_build(const s21_netlist* net, const s21_options* opts, const s21_tran_options* tran,
                          int32_t precision, int64_t* n_ops, int64_t* code_bytes);
/* Host only (no GPU): the AC kernel's generated complex factorisation, for the pattern searched on the
 * matrix at fstart with every Mos1 at a zero small-signal operating point; compiled, not loaded. */
S21_API int s21_jit_build_ac(const s21_netlist* net, const s21_options* opts, double fstart, int32_t precision,
                             int64_t* n_ops, int64_t* code_bytes);
/* Pivot order chosen by the symbolic phase of `group` for its last analysis: internal-to-
 * external row and column maps (AxisMapping.i2e, sparse21/mod.rs:126-162), each [n_vars]. */
S21_API int s21_batch_pivots(const s21_batch* b, int32_t group, int32_t* row_i2e, int32_t* col_i2e);
/* Sizes of the frozen structure: nnz of A, nnz of L+U (incl. fill), LU levels, hub contributions. */
S21_API int s21_batch_structure(const s21_batch* b, int32_t group, int64_t* nnz_a, int64_t* nnz_lu,
                        int64_t* levels, int64_t* n_contrib);
/* Launch shape (no counterpart in the reference, whose solver is single-threaded): instances per
 * wavefront (1, 2, 4, 8 or 16; 0 = chosen per launch from the batch size), and whether the kernel
 * runs a factorisation generated and compiled at run time for the groups' frozen patterns (1 = always,
 * -1 = never, 0 = for a batch of at least 512 instances, or an AC sweep of at least 100,000
 * instance-frequencies).  s21_batch_config reports the last launch: instances per wavefront,
 * wavefronts per workgroup, generated factorisation used.  A batch holding BSIM4 devices always runs the
 * precompiled record-driven factorisation in a kernel that evaluates them (the generated kernels do not). */
S21_API int s21_batch_set_ipw(s21_batch* b, int32_t ipw);
S21_API int s21_batch_set_jit(s21_batch* b, int32_t mode);
/* Launch-shape and structure knobs the autotuner searches (no counterpart in the reference, whose
 * solver is serial): wavefronts per workgroup, at most `waves` (1..8; 0 = the most that still leave a
 * workgroup for every CU), and the hub policy of the stamp scatter (1 = every slot that several
 * components stamp is gathered in order after the device rounds -- one colour round per device class,
 * for latency-bound launches of few instances; 2 = only slots of many contributors, colour rounds for the
 * rest; 0 = chosen by batch size).  The wavefronts change only time; the hub policy also the order in
 * which a shared slot's contributions are added (1 follows the reference's component order exactly; 2
 * adds those of few contributors in colour order, which agrees to rounding). */
S21_API int s21_batch_set_waves(s21_batch* b, int32_t waves);
S21_API int s21_batch_set_hub(s21_batch* b, int32_t policy);
S21_API int s21_batch_config(const s21_batch* b, int32_t* ipw, int32_t* waves, int32_t* jit);
/* AC sweep shape (no counterpart in the reference, which sweeps one instance's frequencies in sequence,
 * each solve starting from the previous frequency's solution, analysis.rs:797-818): each instance's
 * frequencies are cut into `runs` contiguous runs solved in parallel, each run's first frequency
 * started from x = 0 as the reference starts the sweep's first.  The small-signal system does not
 * depend on x, so every frequency converges to the same solution up to rounding; only the Newton
 * iteration counts of the runs' first frequencies differ.  1 = the reference's sequence of starts;
 * 0 (default) = chosen per sweep to fill the device.  s21_batch_last_ac_runs reports the last sweep's. */
S21_API int s21_batch_set_ac_runs(s21_batch* b, int32_t runs);
S21_API int32_t s21_batch_last_ac_runs(const s21_batch* b);
/* Per-phase device cycle counters (s_memtime) of the Newton kernel, for profiling: while on, each
 * instance accumulates [device evaluation, hub gather, residual, factor (+ forward sweep), backward
 * sweep, update, convergence/commit, total] cycles.  Not for batches holding BSIM4 devices (the
 * profiling kernels do not evaluate them: their launches return S21_INVALID). */
S21_API int s21_batch_set_profiling(s21_batch* b, int32_t on);
S21_API int s21_batch_profile(const s21_batch* b, int32_t group, int64_t* out /* [n_inst][8] */);
/* One assembly pass on the GPU (device evaluation + ordered gather) for instance `inst` at the
 * voltages x[n_vars], in OP mode: returns matrix values per created element, in the order
 * Matrix::make created them (row_out/col_out give their external positions), and the RHS
 * (rhs_out[n_vars]).  row_out, col_out and val_out hold `cap` entries; a matrix with more elements
 * returns S21_INVALID and writes nothing. */
S21_API int s21_batch_assemble(s21_batch* b, int32_t group, int32_t inst, const double* x,
                       int32_t* row_out, int32_t* col_out, double* val_out, double* rhs_out,
                       int64_t cap, int64_t* n_elem);

/* ---- single large instance ------------------------------------------------------------------
 * One netlist of tens of thousands of unknowns (a power grid with its loads), whose matrix does not fit a
 * batched instance's LDS region.  The same Newton control, device models, stamps, IC forcing and
 * fixed-step Backward-Euler transient as the batched engine (analysis.rs:169-210, 489-574); the matrix
 * lives in HBM and is factored by the multifrontal method over a fill-reducing order fixed once per
 * analysis (static pivot rows from a maximum transversal, approximate minimum degree) in place of the
 * reference's per-iteration Markowitz search (sparse21/mod.rs:647-991, which grows about as n^2.5).
 * Components: R, C, V, I, Mos1, Diode, BSIM4 (one thread per BSIM4 device evaluates Bsim4::op, tran_op and
 * stamp, bsim4solver.rs:145-3754, tran.rs:11-550, stamp.rs:8-554). */
typedef struct s21_large s21_large;
S21_API int s21_large_create(const s21_netlist* net, const s21_options* opts, int32_t precision, s21_large** out);
S21_API void s21_large_free(s21_large* l);
/* Operating point (analysis.rs:383-388); transient with initial conditions (analysis.rs:489-574). */
S21_API int s21_large_op(s21_large* l);
S21_API int s21_large_tran(s21_large* l, const s21_tran_options* t);
/* Trapezoidal transient (no counterpart in the reference, whose trapezoidal branch panics,
 * analysis.rs:429-435; its companion formula is the one that branch holds: g = 2 C / dt,
 * i = 2 dq / dt - i_prev, each charge with its own previous current).  lte = 0: fixed steps of tstep;
 * lte = 1: steps chosen by the local truncation error of the node voltages (third divided difference over
 * the last four accepted points, in FP64 on the device) against trtol * (reltol * |v| + abstol), at most
 * tstep, a rejected step retried from the last accepted point; results are linearly interpolated onto the
 * reference's output grid (t = 0, tstep, 2 tstep, ...).  "Timestep too small" below dt_min (0: 1e-9 tstep). */
typedef struct {
    int32_t lte;
    double reltol, abstol, trtol, dt_min;
} s21_trap_options;
S21_API int s21_large_tran_trap(s21_large* l, const s21_tran_options* t, const s21_trap_options* o);
/* Accepted and rejected steps of the last trapezoidal transient. */
S21_API int s21_large_trap_stats(const s21_large* l, int64_t* accepted, int64_t* rejected);
/* Unknowns and their names after the last analysis (a transient appends the IC variables). */
S21_API int32_t s21_large_n_vars(const s21_large* l);
S21_API const char* s21_large_var_name(const s21_large* l, int32_t k);
/* Final solution [n_vars]; status (S21_*) and Newton iterations of the last analysis. */
S21_API int s21_large_get_x(const s21_large* l, double* out);
S21_API int s21_large_get_status(const s21_large* l, int32_t* status, int64_t* iters);
/* Transient points stored, their times, and the waveforms of n_sel variables: out[p * n_sel + j]. */
S21_API int64_t s21_large_n_points(const s21_large* l);
S21_API int s21_large_get_axis(const s21_large* l, double* out);
S21_API int s21_large_get_wave(const s21_large* l, int32_t n_sel, const int32_t* vars, double* out);
/* Structure of the last analysis: nnz of A, nnz of L+U, supernodes (fronts), elimination-tree levels,
 * largest front and the factorisation's flops. */
S21_API int s21_large_structure(const s21_large* l, int64_t* nnz_a, int64_t* nnz_lu, int64_t* fronts,
                                int64_t* levels, int64_t* max_front, double* flops_lu);
/* Factor reuse (no counterpart in the reference, which factors at every iteration): 0 factors every
 * Newton iteration; 1 lets an iteration keep the factors of an earlier Jacobian (of an earlier iteration or
 * time point) while the FP64 residual still falls at least fourfold per iteration (at most 32 iterations
 * in a row), so that Newton converges
 * to the same solution under the same test as iterative refinement; a reused solve whose update is not
 * finite is discarded and the iteration repeated with fresh factors.  Default: 1 in S21_MIXED, 0 in
 * S21_FP64.  Not available with BSIM4 devices (default 0, and 1 is rejected): with reused factors a 51-stage
 * BSIM4 ring met non-finite updates, and a three-stage circuit converged to a different trajectory. */
S21_API int s21_large_set_reuse(s21_large* l, int32_t on);
/* Wall time of the last analysis (milliseconds) and factorisations it ran. */
S21_API int s21_large_timing(const s21_large* l, double* total_ms, int64_t* factorisations);
/* Factors the last assembled matrix `reps` times, then sweeps the last residual `reps` times, timed by HIP
 * events on the instance's stream: average milliseconds of each. */
S21_API int s21_large_time_factor(s21_large* l, int32_t reps, double* factor_ms, double* solve_ms);
/* Host only (no GPU): the large path's symbolic phase and the CPU reference of its numeric phase on an
 * n x n matrix of ne entries (duplicates summed): solves A x = b; reports nnz of L+U and flops. */
S21_API int s21_large_host_check(int32_t n, int64_t ne, const int32_t* rows, const int32_t* cols,
                                 const double* vals, const double* b, double relax, double* x,
                                 int64_t* nnz_lu, double* flops_lu);

/* ---- byte ABI mirroring CallableProto::call_bytes (proto.rs:32-145) ------------------------ */
/* Op -> OpResult, Tran -> TranResult, Ac -> AcResult, all protobuf-encoded.  *out is allocated
 * by the library and released with s21_free_bytes.  Op and Tran run the circuit as a batch of one, or
 * on the single-instance path when it holds BSIM4 devices or the batched kernel refuses it as too
 * large; Ac runs as a batch of one only. */
S21_API int s21_op_bytes(const uint8_t* in, size_t len, uint8_t** out, size_t* out_len);
S21_API int s21_tran_bytes(const uint8_t* in, size_t len, uint8_t** out, size_t* out_len);
S21_API int s21_ac_bytes(const uint8_t* in, size_t len, uint8_t** out, size_t* out_len);
S21_API void s21_free_bytes(uint8_t* p);

#ifdef __cplusplus
}
#endif
#endif /* SPICE21_MI355X_H */
