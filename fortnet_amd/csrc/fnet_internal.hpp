// Internal declarations shared by the HIP translation units of libfnet_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <string>
#include <map>
#include <vector>

namespace fnet {

constexpr int kMaxLayers = 8;   // input + hidden + output layers of one element network
constexpr int kMaxDim = 256;    // widest layer of the generic network kernels
constexpr int kMaxMembers = 16; // functions sharing one angular neighbour-pair loop
constexpr int kMaxZ = 16;       // distinct atomic numbers referenced by one model
constexpr int kLadSlots = 10;   // (lambda, 2^m) slots of an exponent-ladder group: 2 x 5
constexpr int kMaxOut = 8;      // network outputs (targets)
constexpr double kPi = 3.14159265358979323846; // lib_dftbp/constants.F90:15

// One radial function (G1/G2/G3), acsf.F90:1287-1373.
// A model whose z-score divides some radial feature by a sigma below |mean| / kRadial64Kappa has
// its radial functions summed in FP64 in the FP32 modes too.  Summed in FP32 over a few dozen
// neighbours they carry a relative error of up to ~4e-7, which the z-score turns into ~1e-5 of a
// network input at this ratio; on a perturbed diamond lattice (|mean| / sigma up to 124) that
// error alone moved the forces by 1.6e-4 eV/A RMS.
constexpr double kRadial64Kappa = 25.0;

struct RadialFn {
  int col;     // output column in G
  int type;    // 1, 2, 3
  int zc;      // element code of neighbour species, -1 = any
  int atomid;  // 0 = none, else 1-based ext column
  double rc, eta, rs, kappa;
};

// Angular functions (G4/G5, acsf.F90:1377-1492) that share type, cutoff, eta, species pair and
// atom-id column: they are evaluated in one pass over the neighbour pairs.
struct AngGroup {
  int type;   // 4 or 5
  int zc1, zc2;
  int same;   // both neighbour lists identical (unresolved or Z1 == Z2): ordered pairs + diagonal
  int atomid;
  int nmem;
  double rc, eta;
  int col[kMaxMembers];
  int xint[kMaxMembers]; // xi as an integer power, or -1 for a general real power
  double xi[kMaxMembers];
  double lam[kMaxMembers];
  double scale[kMaxMembers]; // 2^(1 - xi)
  double pdiag[kMaxMembers]; // (1 + lam)^xi: the j == k terms, where cos(theta) = 1
  // Exponent-ladder form (lad > 0): every member has xi = 2^m with m <= lad <= 4 and one of at
  // most two lambda values; slot 5 i + m stands for (lamv[i], 2^m) and names its member (or -1).
  int lad;
  int lmask;  // bit m set when some member has xi = 2^m
  double lamv[2];
  int smem[kLadSlots];
  double lpdiag[kLadSlots]; // (1 + lamv[i])^(2^m)
  // Adjoint pass: a G5 ladder group and another one over the same lists with the same ladder
  // and lambdas but another eta are evaluated together (partner >= 0 on the first one,
  // follower = 1 on the second).
  int partner;
  int follower;
};

// Network shape of every element network.
struct NetShape {
  int nl;                  // number of layers
  int dims[kMaxLayers];
  long long woff[kMaxLayers]; // offset of ww(l) (d_l x d_{l+1}) in the serialised vector
  long long boff[kMaxLayers]; // offset of bb(l)
  long long nparams;
  int act[64];             // hidden activation per species (<= 64 species)
};

// Per-structure neighbour-search data, computed on the device each step.
struct StructInfo {
  double inv[9];    // inverse lattice, inv[c*3+r] = inv(r,c)
  double smin[3], smax[3]; // extent of fractional coordinates
  double bmin[3], bmax[3]; // Cartesian bounding box of the atoms
  double w[3];      // rc * |row_a(inv L)|: largest fractional offset of a neighbour
  int lo[3], hi[3]; // reference image ranges (latpointiter.F90:226-253)
  int periodic;
  int pad;
};

// Scalar state of one element network's optimizer (fnet_optim.hip); the vectors live in
// fnet_ctx::opt_vecs.
struct OptState {
  int kind, converged, mem, line_search;
  double tol, max_disp, lr;
  // line minimiser (linemin.F90)
  int lm_state, lm_it, lm_conv, lm_maxiter;
  double lm_tol, lm_xx[2], lm_dx[2], lm_xcur, lm_first, lm_maxx;
  int cg_state;        // conjugate gradients: 1 before the first direction
  int lb_it;           // L-BFGS: directions computed so far
  int f_it, f_reset;   // FIRE
  double f_a, f_dt;
};

// Device copy of the model.
struct DevModel {
  int nG = 0, nR = 0, nAG = 0, nZ = 0;
  int maxmem = 1;               // largest angular group
  int all_lad = 0;              // every angular group has the exponent-ladder form
  int has_pairs = 0;            // some angular groups are evaluated in pairs by the adjoint
  int radial_resolved = 0;      // some radial function is restricted to one neighbour element
  int radial64 = 0;             // FP32 modes sum the radial functions in FP64 (kRadial64Kappa)
  int need_ext = 0;             // external feature columns the model reads
  int zlist[kMaxZ];
  int nspecies = 0, nout = 0, nin = 0, next_in = 0;
  double rcmax = 0.0;
  NetShape shape{};
  RadialFn* radial = nullptr;   // [nR]
  AngGroup* groups = nullptr;   // [nAG]
  double* zmean = nullptr;      // [nG] (0 where no standardisation)
  double* zinv = nullptr;       // [nG] (1 where no standardisation)
  int* ext_in = nullptr;        // [next_in] 0-based ext columns
  double* params = nullptr;     // [nspecies][nparams]
  float* params_f = nullptr;    // FP32 copy
  // FP32 weights packed as MFMA operand fragments (per species, stride pk_stride floats):
  // forward fragments of every layer at pk_fw[l], reverse (transposed) fragments at pk_rw[l],
  // biases zero-padded to a multiple of 16 at pk_b[l]
  float* wpack = nullptr;
  long long pk_stride = 0;
  long long pk_fw[kMaxLayers] = {}, pk_rw[kMaxLayers] = {}, pk_b[kMaxLayers] = {};
  // BF16 hi / lo fragments of the hidden-to-hidden layers (BF16 modes), in units of 8 values
  uint16_t* bpack = nullptr;
  long long bk_stride = 0;
  long long bk_fw[kMaxLayers] = {}, bk_rw[kMaxLayers] = {}, bk_sz[kMaxLayers] = {};
  std::vector<double> h_params;
  std::vector<int> h_species_z;
  bool valid = false;
};

// Launch parameters of the batched kernels chosen by the autotuner (fnet_autotune) for a model,
// precision and batch size; 0 / -1 leave the built-in choice.
struct Tune {
  int fwd_wpb = 0;    // waves per block of the forward ACSF kernel (1, 2, 4)
  int adj_wpb = 0;    // waves per block of the adjoint ACSF kernel (1, 2)
  int mlp_waves = 0;  // waves of 16 atoms per block of the MFMA network kernels (1, 2, 4)
  int rbal = -1;      // forward radial lanes in proportion to the element ranges (0, 1)
  int persist = 0;    // ACSF kernels on a persistent grid (1)
  bool operator==(const Tune& o) const {
    return fwd_wpb == o.fwd_wpb && adj_wpb == o.adj_wpb && mlp_waves == o.mlp_waves &&
           rbal == o.rbal && persist == o.persist;
  }
};

struct Timer {
  hipEvent_t ev[8];
  bool made = false;
};

}  // namespace fnet

struct fnet_batch;

struct fnet_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // side stream of the steady-state step (the mirror entries run beside the ACSF forward and
  // the networks) and the events that fork it from and join it to the main stream
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  std::string err;
  fnet::DevModel model;
  hipEvent_t ev[8] = {};
  int n_ev = 0;
  double stage_ms[8] = {};
  double* d_grad = nullptr; // [nspecies][nparams] FP64 gradient of the last train call
  long long grad_cap = 0;
  bool have_grad = false;
  fnet_batch* last_train_batch = nullptr;
  int loss = 0;               // fnet_loss of the training gradient and loss value
  int use_graph = 0;          // fnet_run replays a captured HIP graph of its steady-state step
  double last_loss = 0.0;     // loss of the last training-gradient call
  // device optimizers (fnet_optimizer_init / _step)
  int opt_kind = -1;
  fnet::OptState* opt_states = nullptr;  // [nspecies]
  double* opt_vecs = nullptr;            // [nspecies][opt_vstride]
  int* opt_conv = nullptr;               // [nspecies]
  long long opt_vstride = 0, opt_cap = 0;
  long long model_gen = 0;    // bumped by every model upload (new device tables)
  long long params_gen = 0;   // bumped by every change of the parameters (uploads, optimizer steps)
  double reg_strength = 0.0;  // elastic-net regularisation of the weight gradients
  double reg_alpha = 0.0;
  int n_cu = 0;               // compute units of the device (persistent grids)
  int lds_block = 64 * 1024;  // LDS bytes one block may allocate (the device's limit)
  // autotuned launch parameters per (model generation, precision, log2 of the atom count)
  std::map<std::vector<long long>, fnet::Tune> tunes;
};

struct fnet_batch {
  fnet_ctx* ctx = nullptr;
  int S = 0;
  long long N = 0;
  int next = 0;
  int precision = 1;
  int launch = 0;
  long long capN = 0, capS = 0, capExt = 0;     // allocated capacities (reused by reload)
  long long res_N = 0, res_S = 0;               // capacity of the result buffers
  int res_nG = -1, res_nout = -1;
  std::vector<long long> h_off;
  std::vector<int> h_species;
  // inputs
  double* xyz = nullptr;       // [N][3]
  double* lat = nullptr;       // [S][9]
  int* periodic = nullptr;     // [S]
  long long* off = nullptr;    // [S+1]
  int* atom_struct = nullptr;  // [N]
  int* species = nullptr;      // [N]
  int* zcode = nullptr;        // [N] element slot of each atom under the current model
  std::vector<int> h_atnum;    // atomic numbers (host): element slots follow model uploads
  int zone = 0;                // one element slot, taken by every atom (staging skips the sort)
  unsigned long long zcode_gen = 0;  // model generation zcode was derived with
  double* ext = nullptr;       // [N][next]
  fnet::StructInfo* sinfo = nullptr; // [S]
  // per-species tiles of the network kernels
  int* perm = nullptr;         // [N] atoms grouped by species
  int* tile_sp = nullptr;      // [ntiles]
  int* tile_start = nullptr;   // [ntiles] into perm
  int* tile_cnt = nullptr;     // [ntiles]
  int ntiles = 0;
  int tile_atoms = 0;
  int tile_launch = -1;          // launch mode the tiles were built for
  std::vector<int> h_tile_first; // [S+1] tiles of each structure (per-structure launches)
  // block tiles of the MFMA network kernel (16 atoms per wave, mtile_waves waves per block)
  int* mtile_sp = nullptr;
  int* mtile_start = nullptr;
  int* mtile_cnt = nullptr;
  int nmtiles = 0;
  int mtile_waves = 0;
  std::vector<int> h_mtile_first; // [S+1]
  // neighbour list (CSR)
  int* cnt = nullptr;          // [N]
  int* rowptr = nullptr;       // [N+1]
  int* scan_tmp = nullptr;
  long long scan_tmp_cap = 0;
  int* nj = nullptr;           // [P]
  int* nimg = nullptr;         // [P]
  void* nd = nullptr;          // [P] double4 or float4: image position - centre, and r^2
  long long pcap = 0;
  long long npairs = 0;
  int max_nb = 0;
  int* d_maxnb = nullptr;
  // one-pass search (batched launches): row i at i * nb_cap, then compacted into the CSR list;
  // nb_cap comes from the previous list (0: count pass + fill pass)
  int nb_cap = 0;
  bool nl_padded = false;      // the current list is the padded rows (compacted on download)
  long long pad_len = 0;       // allocated padded entries
  int* pnj = nullptr;
  int* pnimg = nullptr;
  void* pnd = nullptr;         // double4 (room for either precision)
  // cell lists of large periodic structures
  int4* cellinfo = nullptr;    // [S] bins per axis (x, y, z) and use flag (w)
  int* cell_off = nullptr;     // [S+1] first global cell of each structure
  int* cell_count = nullptr;   // [ncells] (reused as the scatter cursor)
  int* cell_start = nullptr;   // [ncells+1]
  int* cell_atoms = nullptr;   // [N]
  int* atom_cell = nullptr;    // [N]
  int4* atom_m = nullptr;      // [N] integer part of the fractional coordinates, local bin
  int* cell_scan_tmp = nullptr;
  long long ncells = 0;
  std::vector<int> h_cellinfo; // [4*S]
  std::vector<double> h_lat;   // [S][9]
  std::vector<int> h_periodic; // [S]
  // results
  void* G = nullptr;           // [N][nG] FP64 (accumulated in FP64 in both precisions)
  void* cG = nullptr;          // [N][nout][nG] T
  double* Eatom = nullptr;     // [N][nout]
  double* Estruct = nullptr;   // [S][nout]
  void* F = nullptr;           // [N][nout][3] FP64
  double* lossgrad = nullptr;  // [N][nout]
  // force assembly (fnet_gather.hip): per-entry adjoint gradients and mirror entries, sized for
  // the list layout (padded rows or CSR)
  double* gvec = nullptr;      // [entries][nout][3], FP64 values (FP64 mode) or FP32 values
  int* mir = nullptr;          // [entries] mirror entry (j, i, -n) of each entry, -1 if none
  int* orph = nullptr;         // [entries] entries without a mirror
  int* d_norph = nullptr;      // number of them
  long long gv_entries = 0;    // allocated entries of mir / orph
  long long gv_len = 0;        // allocated doubles of gvec
  bool have_mirror = false;
  // energy reduction in fixed-size chunks of atoms (two passes, fixed order)
  int* chunk_struct = nullptr; // [nchunks]
  int* chunk_first = nullptr;  // [S+1]
  double* Epart = nullptr;     // [nchunks][nout]
  std::vector<int> h_chunk_first;
  int nchunks = 0;
  long long epart_len = 0;
  // weight gradient: one FP64 partial vector per network tile, reduced in tile order
  double* gpart = nullptr;
  long long gpart_len = 0;
  // steady-state step of fnet_run (no host read-back): the longest row of the list it built is
  // copied to pinned host memory with the step and checked when results are next read
  int* h_maxnb = nullptr;
  bool pending = false;
  long long gen = 0;            // bumped whenever a buffer the step uses is reallocated
  // HIP graph of the steady-state step and the launch parameters it was captured with
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<long long> gkey;
  int graph_failed = 0;         // capture or instantiation failed once: run the step eagerly
  int last_graph = 0;           // the last fnet_run replayed the graph
  double* d_targets = nullptr; // staging of training inputs
  long long targets_cap = 0;
  bool have_nl = false, have_G = false, have_E = false, have_F = false;
  // model and parameter generations the batch's results were computed with
  long long seen_model_gen = -1, seen_params_gen = -1;
  int result_precision = -1;
};

namespace fnet {

// The launch parameters for a batch: the tuned entry of its model, precision and size, else
// the defaults.
const Tune& tune_of(const fnet_batch* b);
std::vector<long long> tune_key(const fnet_batch* b);

// Distance of a neighbour entry.  The neighbour kernels store the squared distance of the FP64
// membership test in V.w, in the list's precision (FP32 rows: rounded to FP32); the ACSF kernels
// take the square root when they stage the entry (entry_r): FP64 rows the FP64 root, FP32 rows
// (the FP32 path, held to its accuracy gates) the FP32 root of the rounded value.  Taking it
// there costs one root per entry instead of one per tested candidate (the compiler evaluates
// the root before the store's branch; ~27 candidates per entry on the Si cells).
__device__ __forceinline__ void entry_d2(double4& v, double d2) { v.w = d2; }
__device__ __forceinline__ void entry_d2(float4& v, double d2) { v.w = (float)d2; }
__device__ __forceinline__ double entry_r(double w) { return sqrt(w); }
__device__ __forceinline__ float entry_r(float w) { return sqrtf(w); }

// launchers (fnet_neigh.hip)
hipError_t launch_struct_prep(fnet_batch* b, double rc, hipStream_t s);
hipError_t launch_neigh_count(fnet_batch* b, double rc, long long a0, long long a1, hipStream_t s);
hipError_t launch_neigh_fill(fnet_batch* b, double rc, long long a0, long long a1, int fp64,
                             hipStream_t s);
// padded rows; a non-null mir (sized for the padded layout) gets -1 at every entry written
hipError_t launch_neigh_pad(fnet_batch* b, double rc, int fp64, hipStream_t s,
                            int* mir = nullptr);
hipError_t launch_cells_pad(fnet_batch* b, double rc, int fp64, hipStream_t s,
                            int* mir = nullptr);
hipError_t launch_compact_rows(fnet_batch* b, int fp64, hipStream_t s);
hipError_t launch_scan(int* cnt, int* rowptr, long long n, int* tmp, hipStream_t s,
                       int* maxout);
// *out = max(cnt[0..n)) (graph-safe: zeroed by a kernel)
hipError_t launch_row_max(const int* cnt, long long n, int* out, hipStream_t s);
long long scan_tmp_size(long long n);

// launchers (fnet_cells.hip)
constexpr int kCellMinAtoms = 64; // structures at least this large may use cell lists
// Cost of one cell-list candidate (bin walk, LDS sort into the reference's order) in units of
// one direct-scan candidate: 8192 Si cells of 290-512 atoms (3 bins of width rc per axis, the
// 27 stencil bins covering the whole cell, ~4 images per atom in the direct scan) took 1.83 ms
// with cell lists and 1.69 ms with the direct scan.
constexpr double kCellCandidateCost = 5.0;
hipError_t launch_cell_bin(fnet_batch* b, hipStream_t s);
// sort capacity of one wave of the cell-list search (power of two >= need, at least 64) and its
// LDS bytes in the writing passes
int cells_sort_capacity(int need);
size_t cells_lds_bytes(int need, int fp64);
hipError_t launch_cells_count(fnet_batch* b, double rc, long long a0, long long a1, hipStream_t s);
hipError_t launch_cells_fill(fnet_batch* b, double rc, long long a0, long long a1, int fp64,
                             hipStream_t s);

// launchers (fnet_optim.hip)
long long opt_vec_doubles(long long n, int mem);
hipError_t launch_opt_init(const double* params, double* vecs, long long n, long long vstride,
                           int mem, int nspecies, hipStream_t s);
hipError_t launch_opt_step(OptState* states, double* params, const double* grad, double* vecs,
                           long long n, long long vstride, double fx, int* conv, int nspecies,
                           hipStream_t s);

// entries of a batch's current neighbour-list layout: padded rows or CSR
inline long long layout_entries(const fnet_batch* b);

// no-allocation force assembly (fnet_gather.hip): static device scratch of kNoallocScratchBytes
// per GPU, its address resolved per device by scratch_init (called by fnet_ctx_create)
constexpr long long kNoallocScratchBytes = 32LL << 20;
hipError_t scratch_init(int device);
void* scratch_ptr(int device);

// launchers (fnet_gather.hip)
hipError_t launch_zero_int(int* p, int n, hipStream_t s);
hipError_t launch_mirror(fnet_batch* b, long long a0, long long a1, long long fill_entries,
                         hipStream_t s);
hipError_t launch_force_gather(fnet_batch* b, int nout, long long a0, long long a1,
                               hipStream_t s);
hipError_t launch_force_orphans(fnet_batch* b, int nout, hipStream_t s);

// launchers (fnet_acsf.hip)
hipError_t launch_acsf_fwd(fnet_batch* b, const DevModel& m, long long a0, long long a1, int fp64,
                           hipStream_t s);
hipError_t launch_acsf_adj(fnet_batch* b, const DevModel& m, long long a0, long long a1, int fp64,
                           hipStream_t s);

// launchers (fnet_mlp.hip)
hipError_t launch_mlp_fwd(fnet_batch* b, const DevModel& m, int first_tile, int ntiles, int fp64,
                          int want_grad_in, hipStream_t s);
constexpr int kEnergyChunk = 2048;  // atoms per block of the energy reduction
hipError_t launch_energy_reduce(fnet_batch* b, const DevModel& m, int s0, int s1, hipStream_t s);
hipError_t launch_lossgrad(fnet_batch* b, const DevModel& m, int n_global, const double* gt,
                           const double* at, const double* aw, const double* dpw, double inv_wsum,
                           int loss, hipStream_t s);
// weight gradient: per-tile partials into b->gpart, then their sum in tile order into grad
hipError_t launch_mlp_grad(fnet_batch* b, const DevModel& m, int fp64, double* grad, hipStream_t s);
// number of weight-gradient partial vectors launch_mlp_grad writes into b->gpart
int grad_parts(const fnet_batch* b, const DevModel& m, int fp64);
// MFMA weight gradient (FP32 / BF16 modes): blocks per MFMA tile, 0 without an MFMA path
int mlp_grad_subtiles(const NetShape& sh, int tile_waves);
hipError_t launch_mlp_grad_mfma(fnet_batch* b, const DevModel& m, hipStream_t s);
// loss into out[0]; scratch: terms [2 S] (per-structure loss and weight), part [3 nchunks]
// (per energy chunk; the chunk tables of the energy reduction must be set up)
hipError_t launch_loss(fnet_batch* b, const DevModel& m, int n_global, const double* gt,
                       const double* at, const double* aw, const double* dpw, int loss,
                       double* out, double* terms, double* part, hipStream_t s);
// elastic-net term (scale = strength / sum of datapoint weights) added to the device gradient
hipError_t launch_regularize(const DevModel& m, double* grad, double scale, double alpha,
                             hipStream_t s);
hipError_t launch_sd_step(double* params, const double* grad, long long n, double lr,
                          double max_disp, double tol, double* scratch, int nspecies,
                          hipStream_t s);
int mlp_tile_atoms(const NetShape& sh, int fp64);
int grad_chunks(int ntiles);  // chunks of the weight-gradient reduction

// launchers (fnet_mlp_mfma.hip): FP32 element networks on MFMA 16x16x4 f32
int mlp_mfma_waves(const NetShape& sh);  // waves per block, 0 if the shape is not supported
void pack_mfma_weights(const NetShape& sh, int nspecies, const double* params,
                       std::vector<float>& out, long long& stride, long long* fw, long long* rw,
                       long long* pb);
void pack_bf16_weights(const NetShape& sh, int nspecies, const double* params,
                       std::vector<uint16_t>& out, long long& stride, long long* bfw,
                       long long* brw, long long* bsz);
hipError_t launch_mlp_mfma(fnet_batch* b, const DevModel& m, int first_tile, int ntiles,
                           int want_grad_in, hipStream_t s);
hipError_t launch_to_float(const double* in, float* out, long long n, hipStream_t s);

size_t lds_bytes_acsf(int max_nb, int nZ, int fp64, int adj);
int acsf_slice_entries(int max_nb);  // entries of one wave's LDS slice for rows up to max_nb

}  // namespace fnet

#define FNET_HIP_TRY(expr)                                   \
  do {                                                       \
    hipError_t _e = (expr);                                  \
    if (_e != hipSuccess) return _e;                         \
  } while (0)

namespace fnet {
inline long long layout_entries(const fnet_batch* b) {
  return b->nl_padded ? (long long)b->N * b->nb_cap : b->npairs;
}
}  // namespace fnet
