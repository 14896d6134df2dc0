/*
 * fnet_hip.h -- C-ABI of the MI355X (gfx950) implementation of Fortnet's BPNN hot path:
 * ACSF descriptors over periodic neighbour lists -> per-element networks -> energies,
 * analytic forces (adjoint, no dense dG/dr) and weight gradients.
 *
 * Plain C types only (ints, doubles, pointers, sizes).  Host arrays are owned by the caller;
 * device memory is owned by the context.  Every entry point returns FNET_OK (0) or a nonzero
 * fnet_status; fnet_last_error() returns the message.  A context is not thread-safe: use one
 * per host thread.  Layouts are row-major with the atom index first, which is the memory order
 * of the reference's column-major Fortran arrays (coords(3,nAtom) == coords[nAtom][3]).
 *
 * The reference has no FFI for this path; it sits behind Fortran type-bound procedures.  Each
 * declaration names the procedure it replaces (vanderhe/fortnet 0.7.4, prog/fortnet/...).
 */
#ifndef FNET_HIP_H
#define FNET_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum fnet_status {
  FNET_OK = 0,
  FNET_ERR_ARG = 1,      /* invalid argument (shape, id, range) */
  FNET_ERR_HIP = 2,      /* HIP runtime error */
  FNET_ERR_STATE = 3,    /* call out of order (no model / no batch / nothing computed) */
  FNET_ERR_CAPACITY = 4, /* a fixed kernel limit was exceeded (message says which) */
};

/* Arithmetic of the device path. */
enum fnet_precision {
  FNET_FP64 = 0, /* everything FP64: the "plain" path, parity ~1e-12 with the FP64 oracle */
  FNET_FP32 = 1, /* ACSF, layers and forces FP32; energy/loss/gradient reductions FP64 */
  FNET_FP32_NET64 = 2, /* ACSF and forces FP32, element networks FP64 (ill-conditioned
                          z-scores: the FP32 input gradient limits the force error) */
  FNET_BF16 = 3,       /* as FNET_FP32, with the hidden-to-hidden layer products on BF16 MFMA
                          with FP32 accumulation, each operand split into bf16 hi + lo parts
                          (hi.hi + hi.lo + lo.hi: within the 1e-5 eV/atom, 1e-4 eV/A gates) */
  FNET_BF16_FAST = 4,  /* as FNET_BF16 with the hi.hi products only (8-bit mantissas; its
                          errors are reported, not within the gates) */
};

/* Launch shape. */
enum fnet_launch {
  FNET_BATCHED = 0,    /* one launch per stage over every structure of the batch */
  FNET_PER_STRUCT = 1, /* one launch per stage per structure (the plain reference shape) */
};

/* G-function types (acsf.F90:40-75) and activations (layer.F90:92-158). */
enum fnet_gtype { FNET_G1 = 1, FNET_G2 = 2, FNET_G3 = 3, FNET_G4 = 4, FNET_G5 = 5 };
enum fnet_activation {
  FNET_TANH = 0, FNET_SIGMOID, FNET_RELU, FNET_LRELU, FNET_SOFTPLUS, FNET_BENT, FNET_ATAN,
  FNET_GAUSSIAN, FNET_HEAVISIDE, FNET_LINEAR
};

typedef struct fnet_ctx fnet_ctx;
typedef struct fnet_batch fnet_batch;

/* Model: the netstat contents (acsf.F90:1908-2135 mapping, bpnn.F90:1141-1288 networks). */
typedef struct fnet_model_desc {
  int n_g;                 /* number of ACSF functions (nG) */
  const int* g_type;       /* [n_g] FNET_G1..FNET_G5 */
  const double* g_rcut;    /* [n_g] cutoff radius, Bohr */
  const double* g_eta;     /* [n_g] */
  const double* g_rs;      /* [n_g] */
  const double* g_kappa;   /* [n_g] */
  const double* g_lambda;  /* [n_g] */
  const double* g_xi;      /* [n_g] */
  const int* g_atomid;     /* [n_g] 0 = none, else 1-based column of ext features */
  const int* g_z1;         /* [n_g] atomic number of neighbour species 1 (0 = unresolved) */
  const int* g_z2;         /* [n_g] atomic number of neighbour species 2 (angular only) */
  const double* z_mean;    /* [n_g] z-score means, or NULL for no standardisation */
  const double* z_std;     /* [n_g] z-score standard deviations, or NULL */
  int n_ext_in;            /* external features appended to the network input (features.F90:227-232) */
  const int* ext_in_idx;   /* [n_ext_in] 1-based columns of ext features */
  int n_species;           /* number of element networks */
  const int* species_z;    /* [n_species] atomic number of each network (BPNN net order) */
  int n_layers;            /* len(dims), input and output layer included */
  const int* dims;         /* [n_layers]; dims[0] == n_g + n_ext_in */
  const int* activation;   /* [n_species] hidden-layer activation (output layer is linear) */
  const double* params;    /* [n_species][n_params] serialised as network.F90:397-427 */
} fnet_model_desc;

/* Batch: packed structures (TGeometry array, typegeometry.F90:24-59). */
typedef struct fnet_batch_desc {
  int n_struct;
  const long long* offsets; /* [n_struct+1] first atom of each structure */
  const double* coords;     /* [n_atoms][3] Cartesian, Bohr, not folded */
  const double* latvecs;    /* [n_struct][3][3], latvecs[s][c][r] = component r of vector c */
  const int* periodic;      /* [n_struct] 0/1 */
  const int* species;       /* [n_atoms] 0-based network index */
  const int* atnum;         /* [n_atoms] atomic numbers (for species-resolved functions) */
  int n_ext;                /* number of external feature columns (may be 0) */
  const double* ext;        /* [n_atoms][n_ext] or NULL */
} fnet_batch_desc;

/* Context: one HIP stream + device buffers.  (No reference counterpart: replaces the process
 * state of prog/fortnet/prg_fnet/fortnet.F90:176-205.) */
int fnet_ctx_create(int device, fnet_ctx** out);
void fnet_ctx_destroy(fnet_ctx* ctx);
int fnet_last_error(const fnet_ctx* ctx, char* buf, int len);
long long fnet_params_per_species(int n_layers, const int* dims); /* network.F90:126-142 */

/* Uploads the ACSF table, z-score and the per-species networks, resident until replaced.
 * Replaces TAcsf%fromFile (acsf.F90:1908-2135) + TBpnn%fromFile (bpnn.F90:1141-1288).  Every
 * batch of the context recomputes its results (neighbour list included) when next read. */
int fnet_model_upload(fnet_ctx* ctx, const fnet_model_desc* model);

/* Replaces only the network parameters (after an optimizer step), same layout as params.  The
 * batches keep their neighbour lists and features and recompute outputs and forces when next
 * read (as after fnet_optimizer_step / fnet_optimizer_step_sd). */
int fnet_params_upload(fnet_ctx* ctx, const double* params);

/* Uploads a batch of structures; the batch stays resident and owns its neighbour state.
 * Replaces the per-structure geometry setup of fnetdata.F90:1141-1162. */
int fnet_batch_upload(fnet_ctx* ctx, const fnet_batch_desc* desc, fnet_batch** out);
void fnet_batch_destroy(fnet_batch* batch);
/* Replaces the structures of a resident batch, reusing its device buffers when they are large
 * enough: the next set of requests of a serving loop.  (No reference counterpart; the i-PI
 * driver of prg_fnet/fortnet.F90:126-169 re-reads one structure per step.) */
int fnet_batch_reload(fnet_batch* batch, const fnet_batch_desc* desc);
/* New coordinates (and lattices, may be NULL to keep) for the same atoms: the serving/MD path. */
int fnet_batch_update_coords(fnet_batch* batch, const double* coords, const double* latvecs);
int fnet_batch_set_mode(fnet_batch* batch, int precision, int launch);
long long fnet_batch_n_atoms(const fnet_batch* batch);

/* Periodic neighbour list with cutoff max(g_rcut), in the reference's entry order and with the
 * reference's FP64 distance test (dynneighlist.F90:233-320, latpointiter.F90:62-253).
 * Replaces the per-(atom, function) TDynNeighList rebuild of acsf.F90:1070-1138.
 * n_pairs receives the total number of entries. */
int fnet_neighbour_list(fnet_ctx* ctx, fnet_batch* batch, long long* n_pairs);
/* Downloads the list: rowptr[n_atoms+1], j[n_pairs] (0-based, global atom index), img[n_pairs]
 * (-1 = central cell, else lexicographic lattice-point index with the origin skipped). */
int fnet_get_neighbours(fnet_ctx* ctx, fnet_batch* batch, long long* rowptr, int* j, int* img);

/* ACSF values (unstandardised) for every atom: TAcsf%calculate -> iGeoAcsf
 * (acsf.F90:540-639,797-866).  G_out [n_atoms][n_g] or NULL to keep them on the device. */
int fnet_acsf_forward(fnet_ctx* ctx, fnet_batch* batch, double* G_out);

/* Network outputs: TBpnn%predictBatch -> iPredict (bpnn.F90:867-900,1001-1058).
 * E_struct [n_struct][n_out] = sum over atoms (FP64 reduction), E_atom [n_atoms][n_out]; either
 * may be NULL.  Runs fnet_acsf_forward first if needed. */
int fnet_energy(fnet_ctx* ctx, fnet_batch* batch, double* E_struct, double* E_atom);

/* Analytic forces F [n_atoms][3*n_out]: forceAnalysis_analytical (forces.F90:317-425) with
 * nJacobian (bpnn.F90:904-997) and calculatePrime (acsf.F90:643-717), computed as one adjoint
 * pass (dE/dG recomputation, never the dense dG/dr).  F may be NULL (keep on device). */
int fnet_forces(fnet_ctx* ctx, fnet_batch* batch, double* F);

/* One full energy+force evaluation on the device (neighbour list, ACSF, networks, adjoint),
 * leaving results on the device; the serving / benchmark step (the predict-with-forces sequence
 * of prg_fnet/fortnet.F90:176-205 for a whole batch).  After the first step of a batch it
 * enqueues every stage without a host read-back: the neighbour search runs with the row capacity
 * of the previous list, and the longest row is checked when results are next read (fnet_energy,
 * fnet_forces, fnet_batch_sync, ...), which recomputes the step if a row outgrew the capacity.
 * It returns before the step has run. */
int fnet_run(fnet_ctx* ctx, fnet_batch* batch);

/* 1: fnet_run replays a HIP graph of its steady-state step, captured on first use and again
 * whenever the batch's buffers, capacities or the model change; 0 (default): eager launches.
 * (No reference counterpart.) */
int fnet_set_graph(fnet_ctx* ctx, int enable);

/* Waits for the batch's last step and validates it (see fnet_run). */
int fnet_batch_sync(fnet_batch* batch);

/* Device time of the whole last fnet_run / fnet_forces step in ms (HIP events on the context
 * stream); *graph = 1 if the batch's last fnet_run replayed its graph. */
int fnet_step_time(fnet_ctx* ctx, double* ms);
int fnet_batch_step_info(const fnet_batch* batch, int* graph);

/* Training gradient of the whole batch: sysTrain + updateGradients + the normalisation of
 * update (bpnn.F90:394-481,610-704,751-767) with the loss gradient set by fnet_set_loss
 * (loss.F90:217-281) and the regularisation set by fnet_set_regularization.
 * g_targets [n_struct][n_global], a_targets [n_atoms][n_out-n_global] (may be NULL when
 * n_out == n_global), atomic_weights [n_atoms] (NULL = 1), dp_weights [n_struct] (NULL = 1).
 * grad [n_species][n_params] receives the FP64 gradient divided by sum(dp_weights);
 * pred [n_atoms][n_out] (may be NULL) the network outputs; loss (may be NULL) the loss value
 * of the data (loss.F90:370-700, without the regularisation term), FP64. */
int fnet_train_gradients(fnet_ctx* ctx, fnet_batch* batch, int n_global, const double* g_targets,
                         const double* a_targets, const double* atomic_weights,
                         const double* dp_weights, double* grad, double* pred,
                         double* loss);

/* Loss of the training gradient and of the reported loss value (TTraining_setLossFunc,
 * initprogram.F90:1195-1227; loss.F90:217-700).  Default FNET_LOSS_MSE. */
typedef enum fnet_loss {
  FNET_LOSS_MSE = 0,
  FNET_LOSS_MAE = 1,
  FNET_LOSS_MAPE = 2,
  FNET_LOSS_RMS = 3
} fnet_loss;
int fnet_set_loss(fnet_ctx* ctx, int loss);

/* Elastic-net regularisation of the weight gradients applied by fnet_train_gradients
 * (the Regularization block, initprogram.F90:1041-1080: Lasso = alpha 1, Ridge = alpha 0;
 * TWeightDerivs_elasticNetRegularization, nestedtypes.F90:336-370, as TBpnn_update applies it,
 * bpnn.F90:750-766).  strength 0 switches it off (default). */
int fnet_set_regularization(fnet_ctx* ctx, double strength, double alpha);

/* Steepest-descent update in FP64 on the device (steepdesc.F90:160-204):
 * x -= lr*g, scaled to max_disp if max|lr*g| > max_disp; converged if max|g| < tol.
 * Acts on the resident parameters and the gradient of the last fnet_train_gradients call;
 * params_out [n_species][n_params] (may be NULL) receives the new parameters. */
int fnet_optimizer_step_sd(fnet_ctx* ctx, double lr, double max_disp, double tol,
                           double* params_out, int* converged);

/* Optimizers of the training loop on the device, FP64, one per element network: the state
 * machines of TOptimizer_next (optimizers.F90:184-221) -- steepest descent (steepdesc.F90:125-204),
 * conjugate gradients (conjgrad.F90), L-BFGS with or without the line minimiser (lbfgs.F90:283-488,
 * linemin.F90) and FIRE (fire.F90) -- with the Training block's settings
 * (initprogram.F90:1130-1185).  fnet_optimizer_init starts them at the resident parameters. */
enum fnet_optimizer { FNET_OPT_SD = 0, FNET_OPT_CG = 1, FNET_OPT_LBFGS = 2, FNET_OPT_FIRE = 3 };
int fnet_optimizer_init(fnet_ctx* ctx, int kind, double tol, double min_disp, double max_disp,
                        double learning_rate, int memory, int line_search);
/* One iteration of every network's optimizer with the gradient (and loss) of the last
 * fnet_train_gradients call: the resident parameters move to the next point to evaluate (as
 * TBpnn_update, bpnn.F90:708-778).  params_out [n_species][n_params] (may be NULL) receives them;
 * converged receives the flag of the last network, as the reference's training loop reads it. */
int fnet_optimizer_step(fnet_ctx* ctx, int kind, double* params_out, int* converged);

/* The force assembly of fnet_forces (adjoint pass, mirror entries of the neighbour list and the
 * per-atom gather) for the batch's resident neighbour list and network input gradients (the
 * batch must have been evaluated with forces in its current mode and with the current
 * parameters: fnet_run or fnet_forces; otherwise FNET_ERR_STATE),
 * writing the batch's force buffer (download with fnet_forces).  No counterpart in the
 * reference: it is for callers that cannot allocate -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  It allocates nothing, does not synchronise
 * and does not copy: its per-entry scratch lives in a static device array of 32 MiB per GPU
 * ((24 n_out + 8) bytes per list entry in FP64, (12 n_out + 8) in the FP32 modes; larger lists
 * are refused with FNET_ERR_CAPACITY).  All calls on one GPU share that array: calls that
 * overlap on one GPU (two streams, two contexts) race. */
int fnet_forces_noalloc(fnet_ctx* ctx, fnet_batch* batch);

/* Free and total device memory in bytes (hipMemGetInfo): the serving loop admits requests by it.
 * (No reference counterpart.) */
int fnet_device_memory(fnet_ctx* ctx, long long* free_bytes, long long* total_bytes);

/* Blocks until the context's stream has drained. */
int fnet_sync(fnet_ctx* ctx);
/* Non-blocking: 0 if all work queued on the context has finished, 1 if some is still running,
 * else an fnet_status (no reference counterpart; used by the serving loop). */
int fnet_query(fnet_ctx* ctx);

/* Timing of the last fnet_run / fnet_forces call, per stage, in milliseconds (HIP events on
 * the context stream): [0] neighbour list (with the mirror entries), [1] ACSF forward,
 * [2] networks, [3] adjoint forces (with the force gather), [4] energy reduction.
 * n = number of entries written (<= 8).  A step replayed as a graph has only its total, in [0]. */
int fnet_stage_times(fnet_ctx* ctx, double* ms, int n);

/* Launch parameters of the batched kernels (no reference counterpart): knobs[5] = waves per
 * block of the forward ACSF kernel (1, 2, 4; 0: default 2), of the adjoint (1, 2; 0: default 1),
 * waves of 16 atoms per block of the MFMA network kernels (1, 2, 4; 0: the most that fit),
 * radial lanes of the forward in proportion to the element ranges (0, 1; -1: when some radial
 * function is species-resolved), ACSF kernels on a persistent grid (0, 1).  They are kept in the
 * context per (model, precision, log2 of the atom count), so batches of about the same size
 * share them; a new model starts from the defaults.  Results do not depend on them, except the
 * radial lane split, which changes the order of the FP64 partial sums of the radial features. */
int fnet_set_tune(fnet_ctx* ctx, fnet_batch* batch, const int* knobs);
int fnet_get_tune(fnet_ctx* ctx, const fnet_batch* batch, int* knobs);
/* Autotuner: times the batch's steady-state step (eager, reps steps per candidate, the least
 * time of the stages a knob changes) over the knobs of fnet_set_tune one after another, keeps
 * the fastest and stores it for the batch's model, precision and size (the defaults when the
 * chosen parameters do not make the whole step faster).  ms[2] (may be NULL)
 * receives the step time with the defaults and with the chosen parameters, knobs[5] the
 * parameters. */
int fnet_autotune(fnet_ctx* ctx, fnet_batch* batch, int reps, double* ms, int* knobs);

#ifdef __cplusplus
}
#endif

#endif /* FNET_HIP_H */
