// Force assembly without atomics: mirror entries of the neighbour list and the per-atom gather.
//
// forceAnalysis_analytical (forces.F90:400-413) contracts dE/dG with dG/dr, where the derivative of
// centre i's functions moves both the centre (+) and each neighbour image (-).  The adjoint pass
// (fnet_acsf.hip) computes g_e = dE/d(image position of entry e) for every entry e = (i, j, n) of
// row i and writes the centre's own sum.  The neighbour side of atom i is the sum of g over the
// entries whose neighbour is i; in a list built with one cutoff those are exactly the mirror
// entries (j, i, -n) of row i's own entries.  The adjoint stores g_e at the slot of e's mirror,
// so every atom gathers -g over its own row's slots, contiguously and in row order: the result
// does not depend on scheduling, and the force array is read and written once per atom instead
// of taking three FP64 atomics per entry (storing g_e at e's own slot and reading it through the
// mirror index gave the same numbers but 5 GB of scattered reads per step on 44 M entries).  The
// FP32 modes store g_e in FP32 (it is the sum of FP32 pair terms, kept in FP64 only while it is
// accumulated); the gather sums in FP64 in both.
//
// The mirror of entry (i, j, n) is found in row j by binary search.  Rows are in the reference's
// order (dynneighlist.F90:233-320): the central cell first, then the lattice points in
// lexicographic order, and within each image the atoms from the centre on, so
// key = (image + 1) * n_struct + (atom - centre) mod n_struct increases strictly along a row.
// Image ranges are symmetric (latpointiter.F90:226-253), so the image index of -n is
// (number of images - 1) - k for image index k (the origin skipped).
//
// The neighbour kernels store this key with each entry (the image index is key / n - 1).
//
// The distance test is evaluated as coords(j) + L n - coords(i) in FP64, whose mirror expression
// rounds differently; a pair exactly at the cutoff can therefore be in one row and not in the
// other.  Such an entry has no mirror: the gather puts it on an orphan list and its -g is added to
// its neighbour's force afterwards (k_force_orphans, the only atomic update left, for pairs within
// an ulp of the cutoff).
#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/fnet_hip.h"
#include "fnet_internal.hpp"

// Scratch of fnet_forces_noalloc: per-entry gradient vectors, mirror entries, the orphan list and
// its counter of one force assembly, for callers that cannot allocate.  Each GPU holds its own
// copy until the process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns on that device (scratch_ptr), never through the array's name.
static __device__ __align__(256) unsigned char scratch[fnet::kNoallocScratchBytes];

namespace fnet {
namespace {

struct RowLayout {
  const int* rowptr;  // CSR rows (cap == 0)
  const int* cnt;     // padded rows: row i at i * cap with cnt[i] entries
  int cap;
  __device__ __forceinline__ long long base(long long i) const {
    return cap ? i * (long long)cap : (long long)rowptr[i];
  }
  __device__ __forceinline__ int len(long long i) const {
    return cap ? min(cnt[i], cap) : rowptr[i + 1] - rowptr[i];
  }
};

__device__ __forceinline__ int mod_n(int x, int n) {
  x %= n;
  return x < 0 ? x + n : x;
}

// One wave per centre; lanes over the centre's entries.  The entries store their keys
// (image + 1) n + (j - i) mod n.  Each mirror pair is searched once, from the entry with i < j (or,
// for the centre's own images, from the one with the lower image index), which writes both
// directions; entries left at -1 (filled before) have no mirror.
__global__ void __launch_bounds__(256)
    k_mirror(RowLayout rl, const int* __restrict__ nj, const int* __restrict__ key,
             const long long* __restrict__ off, const int* __restrict__ atom_struct,
             const StructInfo* __restrict__ sinfo, long long a0, long long a1,
             int* __restrict__ mir) {
  const int lane = threadIdx.x & 63;
  const long long i = a0 + ((long long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  if (i >= a1) return;
  const int s = atom_struct[i];
  const long long o = off[s];
  const int n = (int)(off[s + 1] - o);
  const StructInfo& si = sinfo[s];
  const int nimages =
      si.periodic ? (si.hi[0] - si.lo[0] + 1) * (si.hi[1] - si.lo[1] + 1) *
                            (si.hi[2] - si.lo[2] + 1) - 1
                  : 0;
  const int li = (int)(i - o);
  const long long bi = rl.base(i);
  const int len = rl.len(i);
  for (int e = lane; e < len; e += 64) {
    const int k = key[bi + e];
    const int j = nj[bi + e];
    const int lj = j - (int)o;
    const int img = k / n - 1;
    const int t = k - (img + 1) * n;  // (j - i) mod n
    const int mimg = img < 0 ? -1 : nimages - 1 - img;
    if (!(li < lj || (li == lj && img < mimg))) continue;
    const int target = (mimg + 1) * n + (t == 0 ? 0 : n - t);
    const long long bj = rl.base(j);
    const int lenj = rl.len(j);
    int lo = 0, hi = lenj;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (key[bj + mid] < target) lo = mid + 1;
      else hi = mid;
    }
    if (lo < lenj && key[bj + lo] == target) {
      mir[bi + e] = (int)(bj + lo);
      mir[bj + lo] = (int)(bi + e);
    }
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// F_i += -sum over row i's slots of g (each holding the gradient of the slot's mirror entry), per
// output, in row order, in FP64; entries of row i without a mirror (whose slot holds their own
// gradient) go to the orphan list.  G: the type the adjoint stored g in (its own precision).
template <typename G>
__global__ void __launch_bounds__(256)
    k_force_gather(RowLayout rl, const int* __restrict__ mir, const G* __restrict__ gvec,
                   long long a0, long long a1, int nout, double* __restrict__ F,
                   int* __restrict__ orph, int* __restrict__ norph) {
  const int lane = threadIdx.x & 63;
  const long long i = a0 + ((long long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  if (i >= a1) return;
  const long long bi = rl.base(i);
  const int len = rl.len(i);
  for (int e = lane; e < len; e += 64)
    if (mir[bi + e] < 0) orph[atomicAdd(norph, 1)] = (int)(bi + e);
  for (int t = 0; t < nout; ++t) {
    double x = 0.0, y = 0.0, z = 0.0;
    for (int e = lane; e < len; e += 64) {
      if (mir[bi + e] >= 0) {
        const G* g = gvec + ((bi + e) * nout + t) * 3;
        x -= (double)g[0];
        y -= (double)g[1];
        z -= (double)g[2];
      }
    }
    // x, y, z summed at once: lane bit 5 picks (x, y) or (z, 0), bit 4 one of the two; each
    // addition pairs the same partial sums as a butterfly sum of that component would
    const bool u5 = (lane & 32) != 0, u4 = (lane & 16) != 0;
    double p0 = (u5 ? z : x) + __shfl_xor(u5 ? x : z, 32, 64);
    double p1 = (u5 ? 0.0 : y) + __shfl_xor(u5 ? y : 0.0, 32, 64);
    double q = (u4 ? p1 : p0) + __shfl_xor(u4 ? p0 : p1, 16, 64);
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) q += __shfl_xor(q, m, 64);
    const int c = (u5 ? 2 : 0) + (u4 ? 1 : 0);
    if ((lane & 15) == 0 && c < 3) F[(i * nout + t) * 3 + c] += q;
  }
}

// Entries without a mirror: -g to the neighbour's force.
template <typename G>
__global__ void __launch_bounds__(256)
    k_force_orphans(const int* __restrict__ norph, const int* __restrict__ orph,
                    const int* __restrict__ nj, const G* __restrict__ gvec, int nout,
                    double* __restrict__ F) {
  const int cnt = *norph;
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    const long long e = orph[q];
    const long long j = nj[e];
    for (int t = 0; t < nout; ++t)
      for (int c = 0; c < 3; ++c)
        unsafeAtomicAdd(&F[(j * nout + t) * 3 + c], -(double)gvec[(e * nout + t) * 3 + c]);
  }
}

__global__ void k_zero_int(int* __restrict__ p, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0;
}

__global__ void k_fill_int(int* __restrict__ p, long long n, int v) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

RowLayout layout_of(const fnet_batch* b) {
  RowLayout rl;
  rl.rowptr = b->rowptr;
  rl.cnt = b->cnt;
  rl.cap = b->nl_padded ? b->nb_cap : 0;
  return rl;
}

}  // namespace

hipError_t launch_zero_int(int* p, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_zero_int<<<(n + 255) / 256, 256, 0, s>>>(p, n);
  return hipGetLastError();
}

// Mirror entries of atoms [a0, a1); `entries` > 0 first fills the whole mirror array with -1
// (a range of atoms whose rows are alone in their structures may skip it after the first, and the
// steady-state step skips it: its padded search sets every written entry's slot to -1).
hipError_t launch_mirror(fnet_batch* b, long long a0, long long a1, long long entries,
                         hipStream_t s) {
  if (entries > 0)
    k_fill_int<<<(unsigned)((entries + 255) / 256), 256, 0, s>>>(b->mir, entries, -1);
  const long long n = a1 - a0;
  if (n <= 0) return hipGetLastError();
  const int* nj = b->nl_padded ? b->pnj : b->nj;
  const int* key = b->nl_padded ? b->pnimg : b->nimg;
  k_mirror<<<(unsigned)((n + 3) / 4), 256, 0, s>>>(layout_of(b), nj, key, b->off, b->atom_struct,
                                                   b->sinfo, a0, a1, b->mir);
  return hipGetLastError();
}

hipError_t launch_force_gather(fnet_batch* b, int nout, long long a0, long long a1,
                               hipStream_t s) {
  const long long n = a1 - a0;
  if (n <= 0) return hipSuccess;
  const unsigned grid = (unsigned)((n + 3) / 4);
  if (b->precision == FNET_FP64)
    k_force_gather<double><<<grid, 256, 0, s>>>(layout_of(b), b->mir, b->gvec, a0, a1, nout,
                                                (double*)b->F, b->orph, b->d_norph);
  else
    k_force_gather<float><<<grid, 256, 0, s>>>(layout_of(b), b->mir, (const float*)b->gvec, a0,
                                               a1, nout, (double*)b->F, b->orph, b->d_norph);
  return hipGetLastError();
}

// Per-device address of `scratch`, resolved once per device with that device current.
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = ptr;
  return hipSuccess;
}

void* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

hipError_t launch_force_orphans(fnet_batch* b, int nout, hipStream_t s) {
  const int* nj = b->nl_padded ? b->pnj : b->nj;
  if (b->precision == FNET_FP64)
    k_force_orphans<double><<<1, 256, 0, s>>>(b->d_norph, b->orph, nj, b->gvec, nout,
                                               (double*)b->F);
  else
    k_force_orphans<float><<<1, 256, 0, s>>>(b->d_norph, b->orph, nj, (const float*)b->gvec,
                                              nout, (double*)b->F);
  return hipGetLastError();
}

}  // namespace fnet

using namespace fnet;

namespace {

// The batch's force-assembly buffers pointed at the static scratch for the duration of one
// enqueue (kernel arguments are captured when launched), restored on every exit.
struct ScratchSwap {
  fnet_batch* b;
  double* gvec;
  int *mir, *orph, *norph;
  ScratchSwap(fnet_batch* bb, double* g, int* m, int* o, int* n)
      : b(bb), gvec(bb->gvec), mir(bb->mir), orph(bb->orph), norph(bb->d_norph) {
    b->gvec = g;
    b->mir = m;
    b->orph = o;
    b->d_norph = n;
  }
  ~ScratchSwap() {
    b->gvec = gvec;
    b->mir = mir;
    b->orph = orph;
    b->d_norph = norph;
  }
};

long long align256(long long x) { return (x + 255) & ~255LL; }

int noalloc_fail(fnet_ctx* ctx, int code, const char* msg) {
  ctx->err = msg;
  return code;
}

}  // namespace

extern "C" int fnet_forces_noalloc(fnet_ctx* ctx, fnet_batch* b) {
  if (!ctx) return FNET_ERR_ARG;
  if (!b || b->ctx != ctx) return noalloc_fail(ctx, FNET_ERR_ARG, "batch of another context");
  const DevModel& m = ctx->model;
  if (!m.valid) return noalloc_fail(ctx, FNET_ERR_STATE, "no model uploaded");
  if (!b->have_F || !b->have_nl || b->result_precision != b->precision ||
      b->seen_model_gen != ctx->model_gen || b->seen_params_gen != ctx->params_gen)
    return noalloc_fail(ctx, FNET_ERR_STATE,
                        "fnet_forces_noalloc needs a batch evaluated with forces in its current "
                        "mode and with the current parameters (fnet_run or fnet_forces first)");
  if (hipSetDevice(ctx->device) != hipSuccess)
    return noalloc_fail(ctx, FNET_ERR_HIP, "hipSetDevice failed");
  unsigned char* base = (unsigned char*)scratch_ptr(ctx->device);
  if (!base) return noalloc_fail(ctx, FNET_ERR_STATE, "device scratch not set up");
  const long long entries = std::max<long long>(layout_entries(b), 1);
  const int fp64 = b->precision == FNET_FP64;
  // per-entry gradients in the adjoint's precision (FP32 in the FP32 modes)
  const long long gsize = fp64 ? (long long)sizeof(double) : (long long)sizeof(float);
  const long long gbytes = align256(entries * m.nout * 3 * gsize);
  const long long ibytes = align256(entries * (long long)sizeof(int));
  if (gbytes + 2 * ibytes + 256 > kNoallocScratchBytes)
    return noalloc_fail(ctx, FNET_ERR_CAPACITY,
                        "neighbour list too large for the 32 MiB scratch of "
                        "fnet_forces_noalloc (need (24 n_out + 8) bytes per list entry in FP64, "
                        "(12 n_out + 8) in the FP32 modes)");
  double* gvec = (double*)base;
  int* mir = (int*)(base + gbytes);
  int* orph = (int*)(base + gbytes + ibytes);
  int* norph = (int*)(base + gbytes + 2 * ibytes);
  hipStream_t s = ctx->stream;
  hipError_t e = hipSuccess;
  {
    ScratchSwap sw(b, gvec, mir, orph, norph);
    if (b->launch == FNET_PER_STRUCT) {
      for (int st = 0; st < b->S && e == hipSuccess; ++st)
        e = launch_mirror(b, b->h_off[st], b->h_off[st + 1], st == 0 ? entries : 0, s);
    } else {
      e = launch_mirror(b, 0, b->N, entries, s);
    }
    if (e == hipSuccess) e = launch_zero_int(norph, 1, s);
    if (b->launch == FNET_PER_STRUCT) {
      for (int st = 0; st < b->S && e == hipSuccess; ++st) {
        e = launch_acsf_adj(b, m, b->h_off[st], b->h_off[st + 1], fp64, s);
        if (e == hipSuccess) e = launch_force_gather(b, m.nout, b->h_off[st], b->h_off[st + 1], s);
      }
    } else {
      if (e == hipSuccess) e = launch_acsf_adj(b, m, 0, b->N, fp64, s);
      if (e == hipSuccess) e = launch_force_gather(b, m.nout, 0, b->N, s);
    }
    if (e == hipSuccess) e = launch_force_orphans(b, m.nout, s);
  }
  if (e != hipSuccess) {
    ctx->err = std::string("fnet_forces_noalloc: ") + hipGetErrorString(e);
    return FNET_ERR_HIP;
  }
  b->have_F = true;
  return FNET_OK;
}
