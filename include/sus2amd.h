/*
 * sus2amd.h -- C ABI of the MI355X-native SUS2-SH energy/force/virial (EFS)
 * and loss -> parameter-gradient path.
 *
 * Plain pointers and sizes only.  Every entry point returns SUS2_OK (0) or a
 * status code; sus2_last_error() gives the message of the calling thread's
 * last failure.  "dev" pointers are HIP device memory, "host" pointers are
 * host memory; `stream` is a hipStream_t (NULL = the default stream).
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * superlinear repository):
 *   model handle    MLMTPR(path) / Load / Save / CoeffCount / Coeff
 *                   dev_src/mtpr.h:52, dev_src/mtpr.cpp:4490, :6893;
 *                   interfaces/lammps/ML-SUS2/sus2_sl_model.cpp:174 (Model::load)
 *   EFS             MLMTPR::CalcEFS(Configuration&)  dev_src/mtpr.cpp:8779
 *                   (per-frame energy, forces, PlusStress virial, site energies)
 *   EFS gradient    MLMTPR::AccumulateEFSCombinationGrad  dev_src/mtpr.cpp:9188,
 *                   src/basic_mlip.cpp:421 (ene_weight, frc_weights, str_weights)
 *   loss + gradient NonLinearRegression::CalcObjectiveFunctionGrad
 *                   dev_src/non_linear_regression.cpp:1112 (AddLossGrad :892)
 *   pair compute()  interfaces/lammps/ML-SUS2/sus2_sl_kernel.h:12-43
 *                   (Kernel::begin / forward comm / outer / reverse comm / finish)
 */
#ifndef SUS2AMD_H
#define SUS2AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SUS2_ABI_VERSION 1

enum sus2_status {
  SUS2_OK = 0,
  SUS2_E_INVALID = 1,     /* malformed model or arguments */
  SUS2_E_IO = 2,          /* file cannot be opened / written */
  SUS2_E_UNSUPPORTED = 3, /* a model family or shape without a compiled kernel instance */
  SUS2_E_DEVICE = 4,      /* HIP runtime failure */
  SUS2_E_RANGE = 5        /* species / index / size out of range */
};

enum sus2_precision {
  SUS2_FP64 = 0,  /* all arithmetic in FP64 */
  SUS2_FP32 = 1,  /* FP32 bulk arithmetic, FP64 per-frame energy/virial reductions */
  SUS2_MIXED = 2  /* objective calls only (sus2_loss_grad, sus2_train_lbfgs): FP64 EFS
                     and loss residuals, FP32 parameter-gradient sweeps with FP64 sums */
};

typedef struct sus2_model sus2_model;

int sus2_abi_version(void);
const char* sus2_last_error(void);

/* --- model handle (MLMTPR ctor / Load, Save, CoeffCount, Coeff) ------------ */
int sus2_model_load(const char* path, sus2_model** out);
int sus2_model_save(const sus2_model* m, const char* path);
void sus2_model_free(sus2_model* m);
/* dims[16]: C L K B F M Q R O Qg Kw nbasic nproducts gated species_readout H */
int sus2_model_dims(const sus2_model* m, int* dims);
double sus2_model_cutoff(const sus2_model* m);
int sus2_coeff_count(const sus2_model* m);
/* theta in the reference optimiser layout:
 * shift[C] | scal[2C^2Kw] | radial[F(B+C)] | gate radial[FB] | gate scal |
 * alpha[CF] | gate type[CF] | gate W[F Qg] | gate rank W | outer rank W |
 * species[C] | readout[Q or CQ]   (dev_src/mtpr.cpp:925-975,1586-1740,2108-2210) */
int sus2_coeff_get(const sus2_model* m, double* theta /* host [ncoeff] */);
/* sus2_coeff_set waits for the device's queued work before it replaces the
 * coefficients on the device. */
int sus2_coeff_set(sus2_model* m, const double* theta /* host [ncoeff] */);
/* offsets[13]: shift scal rad grad gscal alpha gtype W gw ow species read ncoeff (-1 = absent) */
int sus2_coeff_offsets(const sus2_model* m, int* offsets);

/* --- model shapes ------------------------------------------------------------
 * No reference counterpart: the reference's kernels take any K and B at run
 * time (interfaces/lammps/ML-SUS2/sus2_sl_model.cpp:276).  Here kernels are
 * compile-time instances per shape (L, K, B = radial basis size, scaling_map
 * mode 0 pair / 1 L / 2 K / 3 LK); shapes without a prebuilt instance are
 * compiled with hipcc at their first evaluation into a cached shape library
 * (SUS2_JIT_DIR, default _jit/ beside the library; SUS2_JIT=0 forbids the
 * compiler run).  sus2_shape_build does that ahead of time (no GPU needed)
 * and writes the shape library's path into path[cap] ("" for a prebuilt
 * shape).  Limits: L <= 6, K (L + 1) <= 64, B <= 16; other shapes fail with
 * SUS2_E_UNSUPPORTED. */
int sus2_shape_build(int L, int K, int B, int scaling_mode, char* path, int cap);

/* --- frame-batch EFS (CalcEFS over a batch; dev pointers) -------------------
 * frame_off [nframes+1] atom offsets (int64), types [natoms] model species,
 * pos [natoms*3] Cartesian, cell [nframes*9] lattice rows, periodic [nframes]
 * (0 = molecule, no images).  Outputs: energy [nframes], force [natoms*3],
 * virial [nframes*6] PlusStress xx yy zz xy xz yz (= -sum d (x) dE/dd),
 * site [natoms] site energies (may be NULL). */
int sus2_efs(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
             const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
             double* energy, double* force, double* virial, double* site, void* stream);

/* --- allocation-free EFS ------------------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as sus2_efs, but every scratch buffer (neighbour lists, per-atom and
 * per-edge state) is carved from one static device array of
 * SUS2_NOALLOC_SCRATCH_BYTES per GPU instead of the handle's growable
 * workspace; the call allocates nothing.  A batch whose scratch exceeds the
 * array fails with SUS2_E_RANGE (the 4096-atom L3/K4 Effect frame fits in
 * both precisions).  The model's device copy must already exist on the
 * current device (made by a first sus2_efs call).  The neighbour counts are
 * read back with plain copies, so the call synchronises `stream`.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams or threads) race and must be serialised by the caller. */
#define SUS2_NOALLOC_SCRATCH_BYTES (128u << 20)
int sus2_efs_noalloc(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                     const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                     double* energy, double* force, double* virial, double* site, void* stream);

/* Same with host buffers (copies in and out; synchronous). */
int sus2_efs_host(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                  const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                  double* energy, double* force, double* virial, double* site);

/* --- parameter gradient (AccumulateEFSCombinationGrad over a batch) ---------
 * grad[ncoeff] += sum_i ene_weight_i dE_i/dtheta
 *               + sum_e w_e . d^2 E / dtheta dd_e,
 * w_e = frc_weight_i - frc_weight_j + str_weight(frame) d_e  for e = (i -> j).
 * ene_weight [natoms], frc_weight [natoms*3], str_weight [nframes*9] row-major,
 * grad [ncoeff] (dev, FP64, accumulated). */
int sus2_efs_grad(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                  const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                  const double* ene_weight, const double* frc_weight, const double* str_weight, double* grad,
                  void* stream);

/* --- objective options (the NonLinearRegression members the trainer sets) ----
 * rel_forces      wgt_rel_forces (train --scale-by-force): when > 0, atom i's
 *                 force weight is wF rel / (|F0_i|^2 + rel)
 * norm_by_forces  n >= 0 (int, the reference's norm_by_forces): a frame's
 *                 energy and force terms are scaled by 0.1 / (0.1 + n mean_i |F0_i|^2)
 * force_loss      SUS2_FORCE_L2 (residual^2) or SUS2_FORCE_LOG_COSH
 *                 (s^2 log cosh(residual / s), s = log_cosh_scale > 0)
 * std_weight, stdd_weight
 *                 per-species site-energy spread regularisers (--std-weight,
 *                 --stdd-weight): std_weight sum_i (E_i - <E>_t)^2
 *                 200 / (200 + |F0_i|^2) / N and stdd_weight
 *                 sum_t (<E>_t - E0 / N)^2 per frame, <E>_t the frame's mean
 *                 site energy of species t
 * constr_weight   legacy radial norm / orthogonality penalty
 *                 constr sum_mu (|c_mu|^2 - 1)^2 + constr sum_{mu<nu} (c_mu . c_nu)^2
 *                 over the outer radial rows (MLMTPR::AddPenaltyGrad), added
 *                 once to the objective; train uses 1e-9 for these models
 * (dev_src/non_linear_regression.cpp:936-1052, dev_src/force_loss.h,
 * dev_src/mtpr.cpp:14050, dev_src/mlp/mtpr_train.cpp:1827-1849,2077-2090,2832).
 * A new handle holds sus2_objective_default: the objective of
 * `mlp-sus2 train` for Static and Effect models (rel_forces 0,
 * norm_by_forces 0, L2, log_cosh_scale 2, no std terms, constr_weight 1e-9).
 * The options apply to sus2_loss_grad, the trainers and graphs created after
 * they are set; sus2_train_linear takes norm_by_forces and constr_weight
 * (the reference's AddToSLAE ignores the others). */
enum { SUS2_FORCE_L2 = 0, SUS2_FORCE_LOG_COSH = 1 };
typedef struct sus2_objective {
  double rel_forces;
  int norm_by_forces;
  int force_loss;
  double log_cosh_scale;
  double std_weight;
  double stdd_weight;
  double constr_weight;
} sus2_objective;
void sus2_objective_default(sus2_objective* out);
int sus2_objective_get(const sus2_model* m, sus2_objective* out);
int sus2_objective_set(sus2_model* m, const sus2_objective* in);

/* --- training objective + gradient (CalcObjectiveFunctionGrad) -------------
 * Labels: e0 [nframes], f0 [natoms*3], s0 [nframes*6] (xx yy zz xy xz yz).
 * weighting: 0 vibrations, 1 structures, 2 molecules; weights wE wF wS.
 * loss_out (host double) = (1/nframes) sum_f L_f + penalty (the options
 * above); grad [ncoeff] (dev FP64) is overwritten with the gradient of
 * loss_out.  Returns once loss_out is known; the gradient kernels finish on
 * `stream` after the return. */
int sus2_loss_grad(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                   const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                   const double* e0, const double* f0, const double* s0, int weighting, double wE, double wF,
                   double wS, double* loss_out, double* grad, void* stream);

/* --- training driver ----------------------------------------------------------
 * L-BFGS (host FP64, `history` pairs; the reference's update rules: Armijo
 * halving with adaptive initial step, Powell-damped curvature pairs, curvature
 * skips, H0 = gamma I) minimising the objective of sus2_loss_grad over one
 * device-resident batch; replaces the L-BFGS loop of MTPR_trainer::Train
 * (dev_src/mtpr_trainer.cpp:4685-4760, src/common/lbfgs.h:114,
 * src/common/lbfgs.cpp:129-327).  The model's coefficients end at the last accepted
 * point.  loss_trace [max_iter+1] (host, may be NULL) receives the objective
 * at the start and after each accepted step; *iters_out the steps taken.
 * Stops when max |grad| <= gtol, after max_iter steps or when the line search
 * fails. */
int sus2_train_lbfgs(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                     const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                     const double* e0, const double* f0, const double* s0, int weighting, double wE, double wF,
                     double wS, int max_iter, int history, double gtol, double* loss_trace, int* iters_out,
                     void* stream);

/* BFGS (host FP64 dense inverse Hessian from the identity, strong-Wolfe line
 * search with c1 = 0.1, c2 = 0.5, the reference's update and curvature guard)
 * over the same objective; replaces BFGS::Iterate as used by the trainer
 * (src/common/bfgs.cpp:251, dev_src/mtpr_trainer.cpp:4685-4760, the default
 * optimizer of example/train-effect.yaml).  loss_trace [max_iter+1] (host,
 * may be NULL) receives the objective at the start and after each accepted
 * step; *iters_out the steps taken.  Stops when max |grad| <= gtol, after
 * max_iter steps or when the line search finds no acceptable step. */
int sus2_train_bfgs(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                    const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                    const double* e0, const double* f0, const double* s0, int weighting, double wE, double wF,
                    double wS, int max_iter, double gtol, double* loss_trace, int* iters_out, void* stream);

/* Linear readout solve: with every other coefficient fixed, the species
 * constants and per-element readout weights enter E, F and the virial
 * linearly; their least-squares fit under the objective's weights (the
 * normal equations of MTPR_trainer::TrainLinear / AddToSLAE,
 * dev_src/mtpr_trainer.cpp:1130-1455, norm_by_forces = 0) is formed on the GPU
 * (design columns from EFS evaluations, J^T J on FP64 MFMA tiles) and solved
 * on the host (Cholesky, `ridge` added to the readout diagonal).  The
 * solution replaces the model's species and readout coefficients;
 * *objective_out (host, may be NULL) receives the objective there (ridge
 * excluded).  Per-element readout models only (else SUS2_E_UNSUPPORTED).
 * Synchronises `stream`. */
int sus2_train_linear(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                      const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                      const double* e0, const double* f0, const double* s0, int weighting, double wE, double wF,
                      double wS, double ridge, double* objective_out, void* stream);

/* --- pair_style sus2/embedded compute() contract ------------------------------
 * Mirrors Kernel (interfaces/lammps/ML-SUS2/sus2_sl_kernel.h:12-43) as driven
 * by PairSUS2Embedded::compute (pair_sus2_embedded.cpp:250-313):
 *   sus2_pair_begin -> forward comm (pack/unpack, width sus2_pair_comm_width)
 *   -> sus2_pair_outer -> reverse comm (pack_reverse/add_reverse)
 *   -> sus2_pair_finish.
 * Atoms [0, nowned) are owned centres, [nowned, nall) ghosts; edges are the
 * owned atoms' full neighbour lists, i owned -> j any, d = x_j - x_i (edges at
 * or beyond the cutoff are dropped).  All buffers are host memory (the LAMMPS
 * atom and comm arrays); computation runs on the GPU.  The comm calls take
 * the atom index lists of pack_forward_comm / unpack_forward_comm /
 * pack_reverse_comm / unpack_reverse_comm with `width` doubles per atom
 * (the gate signal h for Effect models, 0 for Static).  finish writes the
 * energy and virial (xx yy zz xy xz yz) of the owned atoms, site energies of
 * the owned atoms (site may be NULL) and forces on all nall atoms. */
typedef struct sus2_pair sus2_pair;
int sus2_pair_create(sus2_model* m, int precision, sus2_pair** out);
void sus2_pair_free(sus2_pair* p);
int sus2_pair_comm_width(const sus2_pair* p);
int sus2_pair_begin(sus2_pair* p, int nowned, int nall, const int32_t* types /* [nall] */, int64_t nedges,
                    const int32_t* ei, const int32_t* ej, const double* d /* [nedges*3] */);
int sus2_pair_pack_forward(sus2_pair* p, int n, const int32_t* atoms, double* buf /* [n*width] */);
int sus2_pair_unpack_forward(sus2_pair* p, int n, const int32_t* atoms, const double* buf);
int sus2_pair_outer(sus2_pair* p);
int sus2_pair_pack_reverse(sus2_pair* p, int n, const int32_t* atoms, double* buf);
int sus2_pair_add_reverse(sus2_pair* p, int n, const int32_t* atoms, const double* buf);
int sus2_pair_finish(sus2_pair* p, double* energy, double* site /* [nowned] */, double* force /* [nall*3] */,
                     double* virial /* [6] */);

/* --- batches captured in a HIP graph ------------------------------------------
 * One batch's whole EFS (or objective + gradient) call chain captured once
 * and replayed with one hipGraphLaunch: no host waits inside a launch.  The
 * device pointers given at creation are baked in: between launches the
 * caller rewrites positions, cells, species and labels in place (same frame
 * offsets); sus2_coeff_set updates the coefficients the graph reads.  Creation
 * runs the batch three times (host-sized, capture-safe, captured) and
 * synchronises `stream`.  A launch whose neighbour lists outgrow the capacity
 * sized at creation (atoms moved much closer) empties the lists and sets a
 * sticky device flag: its outputs are invalid; sus2_graph_overflow waits for
 * `stream`, returns and clears the flag, and the caller re-creates the graph.
 * The launch shape (sus2_tune_set / sus2_autotune) in force at creation is
 * used.  A graph owns its buffers; do not launch one graph on two streams at
 * once.  No reference counterpart: the reference's batched path is a CPU
 * loop over frames (dev_src/mlp/mtpr_train.cpp:731-790). */
typedef struct sus2_graph sus2_graph;
int sus2_graph_efs_create(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                          const int32_t* types, const double* pos, const double* cell, const int32_t* periodic,
                          double* energy, double* force, double* virial, double* site, void* stream,
                          sus2_graph** out);
/* loss: one device double (the objective); grad [ncoeff] overwritten; energy,
 * force, virial (may be NULL): the batch's EFS outputs. */
int sus2_graph_loss_grad_create(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                                const int32_t* types, const double* pos, const double* cell,
                                const int32_t* periodic, const double* e0, const double* f0, const double* s0,
                                int weighting, double wE, double wF, double wS, double* loss, double* grad,
                                double* energy, double* force, double* virial, void* stream, sus2_graph** out);
int sus2_graph_launch(sus2_graph* g, void* stream);
int sus2_graph_overflow(sus2_graph* g, void* stream, int* overflow);
void sus2_graph_free(sus2_graph* g);

/* --- launch shape of the per-atom EFS kernels and its autotuner -------------
 * The forward passes, edge backwards and gathers run one workgroup per atom
 * (edge backward: per group of eb_atoms = 1, 2 or 4 centres) when grid_mult is
 * 0, else a persistent grid of grid_mult x the resident workgroups striding
 * over the atoms.  Per handle and precision; results agree to rounding
 * (eb_atoms changes the order of per-lane partial sums).  The reference picks
 * its team/vector sizes per kernel at compile time
 * (interfaces/lammps/ML-SUS2-KK/sus2_sl_kokkos_kernel.h:230,277); here
 * sus2_autotune times every candidate (SUS2_TUNE_CANDIDATES: eb_atoms in
 * {1, 2, 4} x grid_mult in {0, 1, 2, 4}, candidate c = (c / 4, c % 4)) on the
 * caller's batch and keeps the fastest. */
#define SUS2_TUNE_CANDIDATES 12
/* Gate-adjoint messages of the batch EFS of Effect models: mode 0 has the
 * outer edge backward write them per edge (E x F values, read by the outer
 * gather); 2 writes them from a messages-only pass after a forces-only one
 * (two lighter kernels instead of one register-bound one); 1 has the outer
 * gather recompute them from the atom's own records and its neighbours'
 * moment adjoints (no per-edge message buffer: 27 GB FP32 at 1.9e8 edges and
 * F = 36); 3 (the default) is automatic: 2 from L = 4 on, where the combined
 * pass is register-bound, else 0; -1 leaves the mode.  *current (may be NULL)
 * receives the mode set.  Modes 0 and 2 give identical results, mode 1 agrees
 * to rounding.  Pair-style stages always use mode 0 (ghost atoms have no
 * records of their own). */
int sus2_tune_messages(sus2_model* m, int precision, int mode, int* current);
int sus2_tune_set(sus2_model* m, int precision, int eb_atoms, int grid_mult);
int sus2_tune_get(sus2_model* m, int precision, int* eb_atoms, int* grid_mult);
int sus2_autotune(sus2_model* m, int precision, int nframes, int natoms, const int64_t* frame_off,
                  const int32_t* types, const double* pos, const double* cell, const int32_t* periodic, int reps,
                  void* stream, int* best /* [2] eb_atoms, grid_mult */, double* ms /* [12] or NULL */);

/* --- per-kernel timing ------------------------------------------------------
 * Replaces the reference's env-gated wall-clock stage timers
 * (SUS2_SH_ACCUM_PROFILE, dev_src/mtpr_sh.cpp:1855-1957; pair_style
 * profile_every, interfaces/lammps/ML-SUS2/pair_sus2_embedded.cpp:320-361)
 * with HIP events attached to every kernel dispatch on its stream.
 * sus2_profile_read waits for pending events and fills up to `cap` entries:
 * names (cap * 64 chars, NUL-terminated), total milliseconds and launch
 * counts; *n receives the number of distinct kernels. */
int sus2_profile_enable(sus2_model* m, int on);
int sus2_profile_reset(sus2_model* m);
int sus2_profile_read(sus2_model* m, int cap, char* names, double* total_ms, int64_t* launches, int* n);
/* Directed-edge count of the last batch's neighbour list (the sum of the
 * reference's Neighborhood::count, src/neighborhoods.h:15, over all atoms). */
int sus2_last_edge_count(const sus2_model* m, int64_t* n);

#ifdef __cplusplus
}
#endif
#endif /* SUS2AMD_H */
