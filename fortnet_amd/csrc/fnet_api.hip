// C-ABI of libfnet_hip.so: context, model and batch management and the stage drivers.
// Declarations and their reference counterparts are in include/fnet_hip.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>

#include "../../include/fnet_hip.h"
#include "fnet_internal.hpp"

using namespace fnet;

namespace fnet {

// Autotuned launch parameters are kept per (model generation, precision, log2 of the atom
// count): a batch of the same model and about the same size reuses them.
std::vector<long long> tune_key(const fnet_batch* b) {
  long long lg = 0;
  while ((1LL << (lg + 1)) <= b->N) ++lg;
  return {b->ctx->model_gen, b->precision, lg};
}

const Tune& tune_of(const fnet_batch* b) {
  static const Tune defaults{};
  if (b->launch != FNET_BATCHED) return defaults;
  const auto it = b->ctx->tunes.find(tune_key(b));
  return it == b->ctx->tunes.end() ? defaults : it->second;
}

}  // namespace fnet

namespace {

int fail(fnet_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int hipfail(fnet_ctx* ctx, hipError_t e, const char* where) {
  std::string m = std::string(where) + ": " + hipGetErrorString(e);
  return fail(ctx, FNET_ERR_HIP, m);
}

#define HIPCHECK(ctx, expr)                                   \
  do {                                                        \
    hipError_t _e = (expr);                                   \
    if (_e != hipSuccess) return hipfail((ctx), _e, #expr);   \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc((void**)p, sizeof(T) * count);
}

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

template <typename T>
hipError_t upload(T** p, const T* h, size_t count, hipStream_t s) {
  hipError_t e = dalloc(p, count);
  if (e != hipSuccess) return e;
  if (count && h) return hipMemcpyAsync(*p, h, sizeof(T) * count, hipMemcpyHostToDevice, s);
  return hipSuccess;
}

void free_model(DevModel& m) {
  dfree(m.radial);
  dfree(m.groups);
  dfree(m.zmean);
  dfree(m.zinv);
  dfree(m.ext_in);
  dfree(m.params);
  dfree(m.params_f);
  dfree(m.wpack);
  m.wpack = nullptr;
  dfree(m.bpack);
  m.bpack = nullptr;
  m.radial = nullptr;
  m.groups = nullptr;
  m.zmean = m.zinv = nullptr;
  m.ext_in = nullptr;
  m.params = nullptr;
  m.params_f = nullptr;
  m.valid = false;
}

// element networks in FP64 (FP64 and FP32_NET64); the others run them in FP32 or BF16
bool net64(int precision) { return precision == FNET_FP64 || precision == FNET_FP32_NET64; }

int zcode_of(const DevModel& m, int z) {
  for (int c = 0; c < m.nZ; ++c)
    if (m.zlist[c] == z) return c;
  return -2;
}

// Fills the exponent-ladder form of an angular group (lad = 0 when the group does not have it):
// every xi a power of two up to 16, at most two lambda values, each (lambda, xi) once.  The usual
// ACSF sets (lambda = +-1, xi = 1, 2, 4, 8, 16) have it.
void ladder_form(AngGroup& a) {
  a.lad = 0;
  for (int s = 0; s < kLadSlots; ++s) {
    a.smem[s] = -1;
    a.lpdiag[s] = 0.0;
  }
  int nl = 0, mx = 1, mask = 0;
  for (int q = 0; q < a.nmem; ++q) {
    int m = -1;
    for (int k = 0; k <= 4; ++k)
      if (a.xint[q] == (1 << k)) m = k;
    if (m < 0) return;
    int i = -1;
    for (int k = 0; k < nl; ++k)
      if (a.lamv[k] == a.lam[q]) i = k;
    if (i < 0) {
      if (nl == 2) return;
      i = nl++;
      a.lamv[i] = a.lam[q];
    }
    if (a.smem[5 * i + m] >= 0) return;
    a.smem[5 * i + m] = q;
    mx = std::max(mx, m);
    mask |= 1 << m;
  }
  if (nl == 0) a.lamv[0] = 0.0;
  if (nl < 2) a.lamv[1] = a.lamv[0];
  for (int i = 0; i < 2; ++i) {
    double p = 1.0 + a.lamv[i];
    for (int m = 0; m <= 4; ++m) {
      a.lpdiag[5 * i + m] = p;
      p = p * p;
    }
  }
  a.lad = mx;
  a.lmask = mask;
}

void invalidate(fnet_batch* b) {
  b->have_nl = b->have_G = b->have_E = b->have_F = false;
  b->have_mirror = false;
  b->pending = false;
}

// Element slots of the batch's atoms under the current model (zcode_of of each atomic number),
// uploaded on the context stream, and whether the model has one slot that every atom takes (the
// ACSF staging then needs no element sort).
int upload_zcode(fnet_batch* b) {
  fnet_ctx* ctx = b->ctx;
  const DevModel& m = ctx->model;
  const size_t N = b->h_atnum.size();
  std::vector<int> zc(N);
  int one = m.nZ == 1;
  for (size_t i = 0; i < N; ++i) {
    zc[i] = zcode_of(m, b->h_atnum[i]);
    one = one && zc[i] == 0;
  }
  b->zone = one;
  b->zcode_gen = ctx->model_gen;
  if (N)
    HIPCHECK(ctx, hipMemcpyAsync(b->zcode, zc.data(), sizeof(int) * N, hipMemcpyHostToDevice,
                                 ctx->stream));
  // the copy reads zc before this function returns
  HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
  return FNET_OK;
}

// Packs the host copy of the parameters into MFMA fragments and uploads them (FP32 networks on
// the matrix cores); a model whose shape the MFMA kernel does not take keeps wpack empty.
int upload_wpack(fnet_ctx* ctx, DevModel& m, hipStream_t s) {
  if (!mlp_mfma_waves(m.shape)) return FNET_OK;
  std::vector<float> pk;
  pack_mfma_weights(m.shape, m.nspecies, m.h_params.data(), pk, m.pk_stride, m.pk_fw, m.pk_rw,
                    m.pk_b);
  if (!m.wpack) HIPCHECK(ctx, dalloc(&m.wpack, pk.size()));
  HIPCHECK(ctx, hipMemcpyAsync(m.wpack, pk.data(), sizeof(float) * pk.size(),
                               hipMemcpyHostToDevice, s));
  // BF16 fragments of the hidden-to-hidden layers (none for networks with one hidden layer)
  std::vector<uint16_t> bk;
  pack_bf16_weights(m.shape, m.nspecies, m.h_params.data(), bk, m.bk_stride, m.bk_fw, m.bk_rw,
                    m.bk_sz);
  if (!bk.empty()) {
    if (!m.bpack) HIPCHECK(ctx, dalloc(&m.bpack, bk.size()));
    HIPCHECK(ctx, hipMemcpyAsync(m.bpack, bk.data(), sizeof(uint16_t) * bk.size(),
                                 hipMemcpyHostToDevice, s));
  }
  HIPCHECK(ctx, hipStreamSynchronize(s));
  return FNET_OK;
}

// (Re)builds the tiles of the network kernels for the batch's launch mode and precision.
int ensure_tiles(fnet_batch* b) {
  fnet_ctx* ctx = b->ctx;
  const DevModel& m = ctx->model;
  const int TA = mlp_tile_atoms(m.shape, net64(b->precision));
  // 16 atoms per wave on the MFMA kernel; the tuned wave count when it fits
  int MW = mlp_mfma_waves(m.shape);
  if (MW && tune_of(b).mlp_waves > 0) MW = std::min(MW, tune_of(b).mlp_waves);
  if (b->tile_atoms == TA && b->mtile_waves == MW && b->tile_launch == b->launch && b->tile_sp)
    return FNET_OK;
  std::vector<int> perm, tsp, tstart, tcnt, msp, mstart, mcnt;
  perm.reserve(b->N);
  b->h_tile_first.assign(b->S + 1, 0);
  b->h_mtile_first.assign(b->S + 1, 0);
  auto add_tiles = [&](int sp, const std::vector<int>& atoms) {
    const int p0 = (int)perm.size();
    for (size_t k = 0; k < atoms.size(); k += TA) {
      tsp.push_back(sp);
      tstart.push_back((int)perm.size());
      const int c = (int)std::min<size_t>(TA, atoms.size() - k);
      tcnt.push_back(c);
      for (int q = 0; q < c; ++q) perm.push_back(atoms[k + q]);
    }
    if (MW)
      for (size_t k = 0; k < atoms.size(); k += 16 * MW) {
        msp.push_back(sp);
        mstart.push_back(p0 + (int)k);
        mcnt.push_back((int)std::min<size_t>(16 * MW, atoms.size() - k));
      }
  };
  if (b->launch == FNET_PER_STRUCT) {
    for (int s = 0; s < b->S; ++s) {
      b->h_tile_first[s] = (int)tsp.size();
      b->h_mtile_first[s] = (int)msp.size();
      for (int sp = 0; sp < m.nspecies; ++sp) {
        std::vector<int> atoms;
        for (long long i = b->h_off[s]; i < b->h_off[s + 1]; ++i)
          if (b->h_species[i] == sp) atoms.push_back((int)i);
        add_tiles(sp, atoms);
      }
    }
    b->h_tile_first[b->S] = (int)tsp.size();
    b->h_mtile_first[b->S] = (int)msp.size();
  } else {
    std::vector<std::vector<int>> by(m.nspecies);
    for (long long i = 0; i < b->N; ++i) by[b->h_species[i]].push_back((int)i);
    for (int sp = 0; sp < m.nspecies; ++sp) add_tiles(sp, by[sp]);
  }
  dfree(b->perm);
  dfree(b->tile_sp);
  dfree(b->tile_start);
  dfree(b->tile_cnt);
  dfree(b->mtile_sp);
  dfree(b->mtile_start);
  dfree(b->mtile_cnt);
  b->perm = b->tile_sp = b->tile_start = b->tile_cnt = nullptr;
  b->mtile_sp = b->mtile_start = b->mtile_cnt = nullptr;
  hipStream_t s = ctx->stream;
  HIPCHECK(ctx, upload(&b->perm, perm.data(), perm.size(), s));
  HIPCHECK(ctx, upload(&b->tile_sp, tsp.data(), tsp.size(), s));
  HIPCHECK(ctx, upload(&b->tile_start, tstart.data(), tstart.size(), s));
  HIPCHECK(ctx, upload(&b->tile_cnt, tcnt.data(), tcnt.size(), s));
  if (MW) {
    HIPCHECK(ctx, upload(&b->mtile_sp, msp.data(), msp.size(), s));
    HIPCHECK(ctx, upload(&b->mtile_start, mstart.data(), mstart.size(), s));
    HIPCHECK(ctx, upload(&b->mtile_cnt, mcnt.data(), mcnt.size(), s));
  }
  HIPCHECK(ctx, hipStreamSynchronize(s));
  b->gen++;
  b->nmtiles = (int)msp.size();
  b->mtile_waves = MW;
  b->ntiles = (int)tsp.size();
  b->tile_atoms = TA;
  b->tile_launch = b->launch;
  return FNET_OK;
}

int ensure_results(fnet_batch* b) {
  fnet_ctx* ctx = b->ctx;
  const DevModel& m = ctx->model;
  if (b->G && b->res_N >= b->N && b->res_S >= b->S && b->res_nG == m.nG && b->res_nout == m.nout)
    return FNET_OK;
  void* pr[] = {b->G, b->cG, b->Eatom, b->Estruct, b->F, b->lossgrad};
  for (void* p : pr) dfree(p);
  const long long N = std::max<long long>(b->N, b->capN), S = std::max<long long>(b->S, b->capS);
  HIPCHECK(ctx, dalloc((double**)&b->G, (size_t)N * m.nG));
  HIPCHECK(ctx, dalloc((double**)&b->cG, (size_t)N * m.nout * m.nG));
  HIPCHECK(ctx, dalloc(&b->Eatom, (size_t)N * m.nout));
  HIPCHECK(ctx, dalloc(&b->Estruct, (size_t)S * m.nout));
  HIPCHECK(ctx, dalloc((double**)&b->F, (size_t)N * m.nout * 3));
  HIPCHECK(ctx, dalloc(&b->lossgrad, (size_t)N * m.nout));
  b->gen++;
  b->res_N = N;
  b->res_S = S;
  b->res_nG = m.nG;
  b->res_nout = m.nout;
  return FNET_OK;
}

void ev_record(fnet_ctx* ctx, int k) {
  if (k < 8) (void)hipEventRecord(ctx->ev[k], ctx->stream);
}

// Room for the one-pass search: nb_cap entries per atom, of either precision.  Returns false
// (and the caller counts first) when the padded rows would take much more memory than the list.
bool ensure_padded(fnet_batch* b) {
  const long long need = (long long)b->N * b->nb_cap;
  if (need <= 0) return false;
  if (need > 4 * std::max<long long>(b->npairs, 1) + (1LL << 22)) return false;
  if (need > b->pad_len) {
    dfree(b->pnj);
    dfree(b->pnimg);
    dfree(b->pnd);
    b->pnj = b->pnimg = nullptr;
    b->pnd = nullptr;
    b->pad_len = 0;
    if (dalloc(&b->pnj, (size_t)need) != hipSuccess ||
        dalloc(&b->pnimg, (size_t)need) != hipSuccess ||
        dalloc((double4**)&b->pnd, (size_t)need) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    b->pad_len = need;
    b->gen++;
  }
  return true;
}

int do_neighbours(fnet_batch* b) {
  fnet_ctx* ctx = b->ctx;
  const DevModel& m = ctx->model;
  hipStream_t s = ctx->stream;
  const double rc = m.rcmax;
  const int fp64 = b->precision == FNET_FP64;
  HIPCHECK(ctx, launch_struct_prep(b, rc, s));
  HIPCHECK(ctx, launch_cell_bin(b, s));
  // One pass into padded rows when the previous list gives a capacity (and the cell-list sort
  // of that capacity fits a block's LDS); a row longer than it sends the batch through the fill
  // pass below, with the counts this pass already made.
  const bool padded = b->launch == FNET_BATCHED && b->nb_cap > 0 &&
                      (b->ncells == 0 ||
                       cells_lds_bytes(b->nb_cap, fp64) <= (size_t)ctx->lds_block) &&
                      ensure_padded(b);
  if (padded) {
    HIPCHECK(ctx, launch_neigh_pad(b, rc, fp64, s));
    HIPCHECK(ctx, launch_cells_pad(b, rc, fp64, s));
  } else if (b->launch == FNET_PER_STRUCT) {
    for (int st = 0; st < b->S; ++st) {
      HIPCHECK(ctx, launch_neigh_count(b, rc, b->h_off[st], b->h_off[st + 1], s));
      HIPCHECK(ctx, launch_cells_count(b, rc, b->h_off[st], b->h_off[st + 1], s));
    }
  } else {
    HIPCHECK(ctx, launch_neigh_count(b, rc, 0, b->N, s));
    HIPCHECK(ctx, launch_cells_count(b, rc, 0, b->N, s));
  }
  HIPCHECK(ctx, launch_scan(b->cnt, b->rowptr, b->N, b->scan_tmp, s, b->d_maxnb));
  int hv[2] = {0, 0};
  HIPCHECK(ctx, hipMemcpyAsync(&hv[0], b->rowptr + b->N, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHECK(ctx, hipMemcpyAsync(&hv[1], b->d_maxnb, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHECK(ctx, hipStreamSynchronize(s));
  const long long P = hv[0];
  if (P < 0) return fail(ctx, FNET_ERR_CAPACITY, "neighbour list exceeds 2^31 entries");
  b->max_nb = hv[1];
  if (b->max_nb > 65535)
    return fail(ctx, FNET_ERR_CAPACITY, "more than 65535 neighbours around one atom");
  if (lds_bytes_acsf(b->max_nb, m.nZ, b->precision == FNET_FP64, 1) > (size_t)ctx->lds_block)
    return fail(ctx, FNET_ERR_CAPACITY,
                "too many neighbours around one atom for the LDS staging (" +
                    std::to_string(b->max_nb) + ")");
  if (P > b->pcap) {
    dfree(b->nj);
    dfree(b->nimg);
    dfree(b->nd);
    b->nj = b->nimg = nullptr;
    b->nd = nullptr;
    const long long cap = P + P / 8 + 1024;
    HIPCHECK(ctx, dalloc(&b->nj, cap));
    HIPCHECK(ctx, dalloc(&b->nimg, cap));
    HIPCHECK(ctx, dalloc((double4**)&b->nd, cap));
    b->pcap = cap;
  }
  b->npairs = P;
  b->nl_padded = padded && b->max_nb <= b->nb_cap;
  if (!b->nl_padded && b->ncells > 0 &&
      cells_lds_bytes(b->max_nb, fp64) > (size_t)ctx->lds_block)
    return fail(ctx, FNET_ERR_CAPACITY,
                "too many neighbours around one atom for the cell-list sort (" +
                    std::to_string(b->max_nb) + ")");
  if (b->nl_padded) {
    // the ACSF kernels read the padded rows; fnet_get_neighbours compacts them
  } else if (b->launch == FNET_PER_STRUCT) {
    for (int st = 0; st < b->S; ++st) {
      HIPCHECK(ctx, launch_neigh_fill(b, rc, b->h_off[st], b->h_off[st + 1], fp64, s));
      HIPCHECK(ctx, launch_cells_fill(b, rc, b->h_off[st], b->h_off[st + 1], fp64, s));
    }
  } else {
    HIPCHECK(ctx, launch_neigh_fill(b, rc, 0, b->N, fp64, s));
    HIPCHECK(ctx, launch_cells_fill(b, rc, 0, b->N, fp64, s));
  }
  if (b->launch == FNET_BATCHED && (b->nb_cap == 0 || b->max_nb > b->nb_cap))
    b->nb_cap = ((b->max_nb + b->max_nb / 4 + 8) + 7) / 8 * 8;
  b->have_nl = true;
  b->have_mirror = false;
  b->result_precision = b->precision;
  return FNET_OK;
}

int do_acsf(fnet_batch* b) {
  fnet_ctx* ctx = b->ctx;
  const DevModel& m = ctx->model;
  const int fp64 = b->precision == FNET_FP64;
  if (b->launch == FNET_PER_STRUCT) {
    for (int st = 0; st < b->S; ++st)
      HIPCHECK(ctx, launch_acsf_fwd(b, m, b->h_off[st], b->h_off[st + 1], fp64, ctx->stream));
  } else {
    HIPCHECK(ctx, launch_acsf_fwd(b, m, 0, b->N, fp64, ctx->stream));
  }
  b->have_G = true;
  return FNET_OK;
}

int do_mlp(fnet_batch* b, int gradin) {
  fnet_ctx* ctx = b->ctx;
  const DevModel& m = ctx->model;
  const int fp64 = net64(b->precision);
  int rc = ensure_tiles(b);
  if (rc) return rc;
  const bool mfma = !fp64 && m.wpack && b->mtile_waves > 0;
  if (b->launch == FNET_PER_STRUCT) {
    for (int st = 0; st < b->S; ++st) {
      if (mfma) {
        const int t0 = b->h_mtile_first[st], t1 = b->h_mtile_first[st + 1];
        HIPCHECK(ctx, launch_mlp_mfma(b, m, t0, t1 - t0, gradin, ctx->stream));
      } else {
        const int t0 = b->h_tile_first[st], t1 = b->h_tile_first[st + 1];
        HIPCHECK(ctx, launch_mlp_fwd(b, m, t0, t1 - t0, fp64, gradin, ctx->stream));
      }
    }
  } else if (mfma) {
    HIPCHECK(ctx, launch_mlp_mfma(b, m, 0, b->nmtiles, gradin, ctx->stream));
  } else {
    HIPCHECK(ctx, launch_mlp_fwd(b, m, 0, b->ntiles, fp64, gradin, ctx->stream));
  }
  return FNET_OK;
}

int do_reduce(fnet_batch* b) {
  fnet_ctx* ctx = b->ctx;
  const DevModel& m = ctx->model;
  if (b->launch == FNET_PER_STRUCT) {
    for (int st = 0; st < b->S; ++st)
      HIPCHECK(ctx, launch_energy_reduce(b, m, st, st + 1, ctx->stream));
  } else {
    HIPCHECK(ctx, launch_energy_reduce(b, m, 0, b->S, ctx->stream));
  }
  b->have_E = true;
  return FNET_OK;
}

// Per-entry adjoint gradients and mirror entries for the current list layout.
int ensure_force_buffers(fnet_batch* b) {
  fnet_ctx* ctx = b->ctx;
  const long long entries = std::max<long long>(layout_entries(b), 1);
  if (entries >= (1LL << 31))
    return fail(ctx, FNET_ERR_CAPACITY, "neighbour list layout exceeds 2^31 entries");
  const long long glen = entries * ctx->model.nout * 3;
  if (entries > b->gv_entries) {
    dfree(b->mir);
    dfree(b->orph);
    b->mir = b->orph = nullptr;
    const long long cap = entries + entries / 8 + 1024;
    HIPCHECK(ctx, dalloc(&b->mir, (size_t)cap));
    HIPCHECK(ctx, dalloc(&b->orph, (size_t)cap));
    b->gv_entries = cap;
    b->gen++;
  }
  if (glen > b->gv_len) {
    dfree(b->gvec);
    b->gvec = nullptr;
    const long long cap = glen + glen / 8 + 4096;
    HIPCHECK(ctx, dalloc(&b->gvec, (size_t)cap));
    b->gv_len = cap;
    b->gen++;
  }
  if (!b->d_norph) HIPCHECK(ctx, dalloc(&b->d_norph, 1));
  return FNET_OK;
}

// Mirror entries of the current list (once per list).
int do_mirror(fnet_batch* b) {
  fnet_ctx* ctx = b->ctx;
  if (b->have_mirror) return FNET_OK;
  int rc = ensure_force_buffers(b);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  const long long entries = layout_entries(b);
  if (b->launch == FNET_PER_STRUCT) {
    for (int st = 0; st < b->S; ++st)
      HIPCHECK(ctx, launch_mirror(b, b->h_off[st], b->h_off[st + 1], st == 0 ? entries : 0, s));
  } else {
    HIPCHECK(ctx, launch_mirror(b, 0, b->N, entries, s));
  }
  b->have_mirror = true;
  return FNET_OK;
}

// Forces: the adjoint writes each entry's gradient and the centre's own sum, then every atom
// gathers its neighbour side through the mirror entries (no atomics, fixed order).
int do_adjoint(fnet_batch* b) {
  fnet_ctx* ctx = b->ctx;
  const DevModel& m = ctx->model;
  const int fp64 = b->precision == FNET_FP64;
  int rc = do_mirror(b);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  HIPCHECK(ctx, launch_zero_int(b->d_norph, 1, s));
  if (b->launch == FNET_PER_STRUCT) {
    for (int st = 0; st < b->S; ++st) {
      HIPCHECK(ctx, launch_acsf_adj(b, m, b->h_off[st], b->h_off[st + 1], fp64, s));
      HIPCHECK(ctx, launch_force_gather(b, m.nout, b->h_off[st], b->h_off[st + 1], s));
    }
  } else {
    HIPCHECK(ctx, launch_acsf_adj(b, m, 0, b->N, fp64, s));
    HIPCHECK(ctx, launch_force_gather(b, m.nout, 0, b->N, s));
  }
  HIPCHECK(ctx, launch_force_orphans(b, m.nout, s));
  b->have_F = true;
  return FNET_OK;
}

// Cell-list set-up of the batch from the lattices: a periodic structure of at least
// kCellMinAtoms atoms whose cell allows >= 3 bins of perpendicular width >= rc on every axis is
// searched with cell lists when that tests fewer candidates, priced at kCellCandidateCost, than
// the direct scan (per centre: 27 bins of n / (nb_0 nb_1 nb_2) atoms, against n atoms in each of
// the ~(1 + 2 w_a) images per axis that the scan's image pruning keeps, w_a = rc / width_a);
// any other structure takes the direct scan.
int setup_cells(fnet_batch* b, const double* lat, const int* per) {
  fnet_ctx* ctx = b->ctx;
  const double rc = ctx->model.rcmax;
  std::vector<int> info(4 * (size_t)b->S, 0), coff(b->S + 1, 0);
  long long total = 0;
  for (int s = 0; s < b->S; ++s) {
    coff[s] = (int)total;
    const long long n = b->h_off[s + 1] - b->h_off[s];
    if (!per[s] || n < kCellMinAtoms) continue;
    const double* L = lat + 9 * s;
    auto o = [&](int r, int c) { return L[c * 3 + r]; };
    const double det = o(0, 0) * (o(1, 1) * o(2, 2) - o(2, 1) * o(1, 2)) -
                       o(0, 1) * (o(1, 0) * o(2, 2) - o(2, 0) * o(1, 2)) +
                       o(0, 2) * (o(1, 0) * o(2, 1) - o(2, 0) * o(1, 1));
    int nb[3];
    bool ok = true;
    double direct = 1.0;  // images per atom that the direct scan tests
    for (int a = 0; a < 3; ++a) {
      // row a of inv(L) = (cofactor column a) / det
      const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
      double ss = 0.0;
      for (int c = 0; c < 3; ++c) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        const double cof = o(c1, a1) * o(c2, a2) - o(c1, a2) * o(c2, a1);
        ss += (cof / det) * (cof / det);
      }
      const double w = rc * std::sqrt(ss);
      const double nbf = std::floor(1.0 / (w * (1.0 + 1e-9) + 1e-12));
      nb[a] = nbf > 1000.0 ? 1000 : (int)nbf;
      ok = ok && nb[a] >= 3;
      direct *= 1.0 + 2.0 * w;
    }
    if (!ok) continue;
    const double cells = kCellCandidateCost * 27.0 / ((double)nb[0] * nb[1] * nb[2]);
    if (cells >= direct) continue;
    info[4 * s + 0] = nb[0];
    info[4 * s + 1] = nb[1];
    info[4 * s + 2] = nb[2];
    info[4 * s + 3] = 1;
    total += (long long)nb[0] * nb[1] * nb[2];
  }
  coff[b->S] = (int)total;
  if (total >= (1LL << 30)) return fail(ctx, FNET_ERR_CAPACITY, "too many cells");
  hipStream_t s = ctx->stream;
  if (!b->cellinfo) {
    HIPCHECK(ctx, dalloc(&b->cellinfo, (size_t)b->S));
    HIPCHECK(ctx, dalloc(&b->cell_off, (size_t)b->S + 1));
    HIPCHECK(ctx, dalloc(&b->atom_cell, (size_t)b->N));
    HIPCHECK(ctx, dalloc(&b->atom_m, (size_t)b->N));
    HIPCHECK(ctx, dalloc(&b->cell_atoms, (size_t)b->N));
  }
  if (total != b->ncells) {
    dfree(b->cell_count);
    dfree(b->cell_start);
    dfree(b->cell_scan_tmp);
    b->cell_count = b->cell_start = b->cell_scan_tmp = nullptr;
    HIPCHECK(ctx, dalloc(&b->cell_count, (size_t)total));
    HIPCHECK(ctx, dalloc(&b->cell_start, (size_t)total + 1));
    HIPCHECK(ctx, dalloc(&b->cell_scan_tmp, (size_t)scan_tmp_size(total)));
    b->ncells = total;
    b->gen++;
  }
  b->h_cellinfo = info;
  HIPCHECK(ctx, hipMemcpyAsync(b->cellinfo, info.data(), sizeof(int) * info.size(),
                               hipMemcpyHostToDevice, s));
  HIPCHECK(ctx, hipMemcpyAsync(b->cell_off, coff.data(), sizeof(int) * coff.size(),
                               hipMemcpyHostToDevice, s));
  HIPCHECK(ctx, hipStreamSynchronize(s));
  return FNET_OK;
}

// The full step with the host in the loop: the neighbour search reads the list size and the
// longest row back before the ACSF stages are sized (the first step of a batch, the per-structure
// launches, and the redo of a steady-state step whose list outgrew its capacities).
int run_sync_step(fnet_batch* b) {
  fnet_ctx* ctx = b->ctx;
  invalidate(b);
  b->last_graph = 0;
  ev_record(ctx, 0);
  int rc = do_neighbours(b);
  if (rc) return rc;
  ev_record(ctx, 1);
  rc = do_acsf(b);
  if (rc) return rc;
  ev_record(ctx, 2);
  rc = do_mlp(b, 1);
  if (rc) return rc;
  ev_record(ctx, 3);
  rc = do_adjoint(b);
  if (rc) return rc;
  ev_record(ctx, 4);
  rc = do_reduce(b);
  if (rc) return rc;
  ev_record(ctx, 5);
  ctx->n_ev = 6;
  return FNET_OK;
}

// A batched step can skip every host read-back once a previous list gave the row capacity of the
// one-pass search (nb_cap) and the longest row the ACSF slices are sized for (max_nb); buffers
// are sized here, outside any capture.
bool fast_step_ready(fnet_batch* b) {
  fnet_ctx* ctx = b->ctx;
  const DevModel& m = ctx->model;
  if (b->launch != FNET_BATCHED || b->N <= 0 || b->nb_cap <= 0 || b->max_nb <= 0) return false;
  if (acsf_slice_entries(b->max_nb) > b->nb_cap) return false;
  if (lds_bytes_acsf(b->max_nb, m.nZ, b->precision == FNET_FP64, 1) > (size_t)ctx->lds_block)
    return false;
  if (b->ncells > 0 &&
      cells_lds_bytes(b->nb_cap, b->precision == FNET_FP64) > (size_t)ctx->lds_block)
    return false;
  if (!ensure_padded(b)) return false;
  const bool was = b->nl_padded;
  b->nl_padded = true;
  if (ensure_force_buffers(b)) {
    b->nl_padded = was;
    (void)hipGetLastError();
    return false;
  }
  if (!b->h_maxnb) {
    if (hipHostMalloc((void**)&b->h_maxnb, sizeof(int), hipHostMallocDefault) != hipSuccess) {
      b->h_maxnb = nullptr;
      (void)hipGetLastError();
      return false;
    }
    *b->h_maxnb = 0;
  }
  return true;
}

// Kernels of one steady-state step: no host read-back, no allocation and no memset, so the
// sequence runs eagerly or as a captured HIP graph.  The padded rows are searched with capacity
// nb_cap and every row is cut to the ACSF slice of the previous longest row; the longest row of
// this list is reduced on the device and copied to pinned memory with the step (validate_step).
int enqueue_fast_step(fnet_batch* b, bool events) {
  fnet_ctx* ctx = b->ctx;
  const DevModel& m = ctx->model;
  hipStream_t s = ctx->stream;
  const double rc = m.rcmax;
  const int fp64 = b->precision == FNET_FP64;
  if (events) ev_record(ctx, 0);
  HIPCHECK(ctx, launch_struct_prep(b, rc, s));
  HIPCHECK(ctx, launch_cell_bin(b, s));
  // the search clears the mirror slot of every entry it writes, so the mirror search needs no fill
  // of the whole padded array first
  HIPCHECK(ctx, launch_neigh_pad(b, rc, fp64, s, b->mir));
  HIPCHECK(ctx, launch_cells_pad(b, rc, fp64, s, b->mir));
  HIPCHECK(ctx, launch_row_max(b->cnt, b->N, b->d_maxnb, s));
  // The mirror entries need only the list: their search (latency-bound) runs on the side stream
  // beside the ACSF forward and the networks (VALU- and matrix-bound), and the adjoint waits for
  // it.  Under graph capture the event pair becomes a fork and a join of the graph.
  HIPCHECK(ctx, hipEventRecord(ctx->fork_ev, s));
  HIPCHECK(ctx, hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0));
  HIPCHECK(ctx, launch_mirror(b, 0, b->N, 0, ctx->side));
  HIPCHECK(ctx, hipEventRecord(ctx->join_ev, ctx->side));
  if (events) ev_record(ctx, 1);
  HIPCHECK(ctx, launch_acsf_fwd(b, m, 0, b->N, fp64, s));
  if (events) ev_record(ctx, 2);
  const bool mfma = !net64(b->precision) && m.wpack && b->mtile_waves > 0;
  if (mfma) HIPCHECK(ctx, launch_mlp_mfma(b, m, 0, b->nmtiles, 1, s));
  else HIPCHECK(ctx, launch_mlp_fwd(b, m, 0, b->ntiles, net64(b->precision), 1, s));
  if (events) ev_record(ctx, 3);
  HIPCHECK(ctx, hipStreamWaitEvent(s, ctx->join_ev, 0));
  HIPCHECK(ctx, launch_acsf_adj(b, m, 0, b->N, fp64, s));
  HIPCHECK(ctx, launch_zero_int(b->d_norph, 1, s));
  HIPCHECK(ctx, launch_force_gather(b, m.nout, 0, b->N, s));
  HIPCHECK(ctx, launch_force_orphans(b, m.nout, s));
  if (events) ev_record(ctx, 4);
  HIPCHECK(ctx, launch_energy_reduce(b, m, 0, b->S, s));
  if (events) ev_record(ctx, 5);
  HIPCHECK(ctx, hipMemcpyAsync(b->h_maxnb, b->d_maxnb, sizeof(int), hipMemcpyDeviceToHost, s));
  return FNET_OK;
}

void drop_graph(fnet_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  b->gkey.clear();
}

// What the captured kernels' launch parameters depend on.
std::vector<long long> graph_key(const fnet_batch* b) {
  const fnet_ctx* ctx = b->ctx;
  const Tune& t = tune_of(b);
  return {b->gen, ctx->model_gen, b->precision, b->launch, b->nb_cap, b->max_nb, b->ncells,
          b->N, b->S, b->ntiles, b->nmtiles, t.fwd_wpb, t.adj_wpb, t.mlp_waves, t.rbal,
          t.persist};
}

int run_fast_step(fnet_batch* b) {
  fnet_ctx* ctx = b->ctx;
  hipStream_t s = ctx->stream;
  b->have_nl = b->have_G = b->have_E = b->have_F = false;
  b->nl_padded = true;
  b->last_graph = 0;
  if (ctx->use_graph && !b->graph_failed) {
    const std::vector<long long> key = graph_key(b);
    if (!b->gexec || key != b->gkey) {
      drop_graph(b);
      HIPCHECK(ctx, hipStreamSynchronize(s));
      hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
      int rc = e == hipSuccess ? enqueue_fast_step(b, false) : FNET_ERR_HIP;
      hipGraph_t g = nullptr;
      const hipError_t e2 = hipStreamEndCapture(s, &g);
      if (e == hipSuccess && rc == FNET_OK && e2 == hipSuccess && g) {
        b->graph = g;
        if (hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0) != hipSuccess) b->gexec = nullptr;
      } else if (g) {
        (void)hipGraphDestroy(g);
      }
      if (!b->gexec) {
        // no graph for this batch: its steps run eagerly from here on
        drop_graph(b);
        (void)hipGetLastError();
        b->graph_failed = 1;
      } else {
        b->gkey = key;
      }
    }
    if (b->gexec) {
      ev_record(ctx, 0);
      HIPCHECK(ctx, hipGraphLaunch(b->gexec, s));
      ev_record(ctx, 1);
      ctx->n_ev = 2;
      b->last_graph = 1;
    }
  }
  if (!b->last_graph) {
    const int rc = enqueue_fast_step(b, true);
    if (rc) return rc;
    ctx->n_ev = 6;
  }
  b->have_nl = b->have_G = b->have_E = b->have_F = true;
  b->have_mirror = true;
  b->result_precision = b->precision;
  b->pending = true;
  return FNET_OK;
}

// Checks the longest row of a steady-state step once it has run.  A row longer than the ACSF
// slice (or the row capacity) was cut, so that step is recomputed on the synchronous path, which
// also raises the capacities for the next steps; a list that shrank well below its slice lowers
// the slice.
int validate_step(fnet_batch* b) {
  if (!b->pending) return FNET_OK;
  fnet_ctx* ctx = b->ctx;
  HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
  b->pending = false;
  const int mx = *b->h_maxnb;
  if (mx <= acsf_slice_entries(b->max_nb) && mx <= b->nb_cap) {
    if (mx + 16 < b->max_nb) b->max_nb = mx;
    return FNET_OK;
  }
  return run_sync_step(b);
}

int check_ready(fnet_ctx* ctx, fnet_batch* b) {
  if (!ctx) return FNET_ERR_ARG;
  if (!b) return fail(ctx, FNET_ERR_ARG, "batch is NULL");
  if (b->ctx != ctx) return fail(ctx, FNET_ERR_ARG, "batch belongs to another context");
  if (!ctx->model.valid) return fail(ctx, FNET_ERR_STATE, "no model uploaded");
  HIPCHECK(ctx, hipSetDevice(ctx->device));
  // results of an earlier model (list cutoff, functions, networks) or of earlier parameters
  // are stale
  if (b->seen_model_gen != ctx->model_gen) {
    invalidate(b);
    b->seen_model_gen = ctx->model_gen;
    if (b->zcode_gen != ctx->model_gen) {
      const int rc = upload_zcode(b);
      if (rc) return rc;
    }
  } else if (b->seen_params_gen != ctx->params_gen) {
    b->have_E = b->have_F = false;
  }
  b->seen_params_gen = ctx->params_gen;
  return ensure_results(b);
}

int download_T(fnet_ctx* ctx, const void* dsrc, int fp64, size_t n, double* out) {
  if (fp64) {
    HIPCHECK(ctx, hipMemcpyAsync(out, dsrc, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
  } else {
    std::vector<float> tmp(n);
    HIPCHECK(ctx, hipMemcpyAsync(tmp.data(), dsrc, sizeof(float) * n, hipMemcpyDeviceToHost,
                                 ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t k = 0; k < n; ++k) out[k] = (double)tmp[k];
  }
  return FNET_OK;
}

}  // namespace

extern "C" {

int fnet_ctx_create(int device, fnet_ctx** out) {
  if (!out) return FNET_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FNET_ERR_HIP;
  if (device < 0 || device >= ndev) return FNET_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return FNET_ERR_HIP;
  // the device's copy of the no-allocation scratch: its address is taken here, once per device
  if (scratch_init(device) != hipSuccess) return FNET_ERR_HIP;
  fnet_ctx* ctx = new fnet_ctx();
  ctx->device = device;
  if (hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, device) !=
      hipSuccess)
    ctx->n_cu = 0;
  // a block's LDS: 160 KiB on gfx950 (the whole CU's), which the per-centre staging of the ACSF
  // kernels may take when a structure's rows are long (one wave per block then)
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) ==
          hipSuccess &&
      lds > 0)
    ctx->lds_block = lds;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return FNET_ERR_HIP;
  }
  for (int k = 0; k < 8; ++k) (void)hipEventCreate(&ctx->ev[k]);
  if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming) != hipSuccess) {
    fnet_ctx_destroy(ctx);
    return FNET_ERR_HIP;
  }
  *out = ctx;
  return FNET_OK;
}

void fnet_ctx_destroy(fnet_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  free_model(ctx->model);
  dfree(ctx->d_grad);
  dfree(ctx->opt_vecs);
  dfree(ctx->opt_states);
  dfree(ctx->opt_conv);
  for (int k = 0; k < 8; ++k) (void)hipEventDestroy(ctx->ev[k]);
  if (ctx->side) {
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamDestroy(ctx->side);
  }
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int fnet_last_error(const fnet_ctx* ctx, char* buf, int len) {
  if (!ctx || !buf || len <= 0) return FNET_ERR_ARG;
  std::snprintf(buf, (size_t)len, "%s", ctx->err.c_str());
  return FNET_OK;
}

long long fnet_params_per_species(int n_layers, const int* dims) {
  if (n_layers < 2 || !dims) return -1;
  long long nw = 0, nb = 0;
  for (int l = 0; l < n_layers - 1; ++l) nw += (long long)dims[l] * dims[l + 1];
  nw += dims[n_layers - 1];
  for (int l = 0; l < n_layers; ++l) nb += dims[l];
  return nw + nb;
}

int fnet_model_upload(fnet_ctx* ctx, const fnet_model_desc* md) {
  if (!ctx) return FNET_ERR_ARG;
  if (!md) return fail(ctx, FNET_ERR_ARG, "model descriptor is NULL");
  HIPCHECK(ctx, hipSetDevice(ctx->device));
  HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
  const int nG = md->n_g;
  if (nG < 0 || (nG > 0 && (!md->g_type || !md->g_rcut)))
    return fail(ctx, FNET_ERR_ARG, "invalid ACSF table");
  if (md->n_species <= 0 || md->n_species > 64)
    return fail(ctx, FNET_ERR_ARG, "n_species must be in 1..64");
  if (md->n_layers < 2 || md->n_layers > kMaxLayers)
    return fail(ctx, FNET_ERR_ARG, "n_layers must be in 2..8");
  if (!md->dims || !md->params || !md->activation || !md->species_z)
    return fail(ctx, FNET_ERR_ARG, "network arrays missing");
  if (md->dims[0] != nG + md->n_ext_in)
    return fail(ctx, FNET_ERR_ARG, "dims[0] must equal n_g + n_ext_in");
  for (int l = 0; l < md->n_layers; ++l)
    if (md->dims[l] <= 0 || md->dims[l] > kMaxDim)
      return fail(ctx, FNET_ERR_ARG, "layer width must be in 1..256");
  if (md->dims[md->n_layers - 1] > kMaxOut)
    return fail(ctx, FNET_ERR_ARG, "at most 8 network outputs");
  DevModel m;
  m.nG = nG;
  m.nspecies = md->n_species;
  m.nout = md->dims[md->n_layers - 1];
  m.nin = md->dims[0];
  m.next_in = md->n_ext_in;
  // element table
  std::vector<int> zs;
  for (int sp = 0; sp < md->n_species; ++sp) zs.push_back(md->species_z[sp]);
  for (int g = 0; g < nG; ++g) {
    if (md->g_z1 && md->g_z1[g]) zs.push_back(md->g_z1[g]);
    if (md->g_z2 && md->g_z2[g]) zs.push_back(md->g_z2[g]);
  }
  std::sort(zs.begin(), zs.end());
  zs.erase(std::unique(zs.begin(), zs.end()), zs.end());
  if ((int)zs.size() > kMaxZ) return fail(ctx, FNET_ERR_CAPACITY, "more than 16 elements");
  m.nZ = (int)zs.size();
  for (int c = 0; c < m.nZ; ++c) m.zlist[c] = zs[c];
  // functions
  std::vector<RadialFn> rad;
  std::vector<AngGroup> grp;
  std::map<std::tuple<int, double, double, int, int, int>, int> open;
  std::vector<double> zmean(nG, 0.0), zinv(nG, 1.0);
  double rcmax = 0.0;
  for (int g = 0; g < nG; ++g) {
    const int ty = md->g_type[g];
    const double rc = md->g_rcut[g];
    if (!(rc > 0.0)) return fail(ctx, FNET_ERR_ARG, "cutoff must be positive");
    rcmax = std::max(rcmax, rc);
    const int z1 = md->g_z1 ? md->g_z1[g] : 0, z2 = md->g_z2 ? md->g_z2[g] : 0;
    const int atomid = md->g_atomid ? md->g_atomid[g] : 0;
    const double eta = md->g_eta ? md->g_eta[g] : 0.0;
    if (atomid < 0) return fail(ctx, FNET_ERR_ARG, "negative atom id");
    if (ty >= 1 && ty <= 3) {
      RadialFn f{};
      f.col = g;
      f.type = ty;
      f.zc = z1 ? zcode_of(m, z1) : -1;
      f.atomid = atomid;
      f.rc = rc;
      f.eta = eta;
      f.rs = md->g_rs ? md->g_rs[g] : 0.0;
      f.kappa = md->g_kappa ? md->g_kappa[g] : 0.0;
      rad.push_back(f);
    } else if (ty == 4 || ty == 5) {
      if ((z1 == 0) != (z2 == 0))
        return fail(ctx, FNET_ERR_ARG, "angular function with one species of two unresolved");
      const auto key = std::make_tuple(ty, rc, eta, z1, z2, atomid);
      auto it = open.find(key);
      if (it == open.end() || grp[it->second].nmem == kMaxMembers) {
        AngGroup a{};
        a.type = ty;
        a.zc1 = z1 ? zcode_of(m, z1) : -1;
        a.zc2 = z2 ? zcode_of(m, z2) : -1;
        a.same = (z1 == z2) ? 1 : 0;
        a.atomid = atomid;
        a.rc = rc;
        a.eta = eta;
        grp.push_back(a);
        open[key] = (int)grp.size() - 1;
        it = open.find(key);
      }
      AngGroup& a = grp[it->second];
      const double xi = md->g_xi ? md->g_xi[g] : 0.0;
      const int q = a.nmem++;
      a.col[q] = g;
      a.xi[q] = xi;
      a.lam[q] = md->g_lambda ? md->g_lambda[g] : 0.0;
      a.xint[q] = (xi >= 0.0 && xi <= 64.0 && std::floor(xi) == xi) ? (int)xi : -1;
      a.scale[q] = std::pow(2.0, 1.0 - xi);
      a.pdiag[q] = std::pow(1.0 + a.lam[q], xi);
    } else {
      return fail(ctx, FNET_ERR_ARG, "invalid G-function type");
    }
    if (md->z_mean && md->z_std && !(md->z_std[g] < 1e-8)) {
      zmean[g] = md->z_mean[g];
      zinv[g] = 1.0 / md->z_std[g];
    }
  }
  m.nR = (int)rad.size();
  m.radial64 = 0;
  for (const RadialFn& f : rad)
    if (std::fabs(zmean[f.col]) * zinv[f.col] > kRadial64Kappa)
      m.radial64 = 1;
  m.radial_resolved = 0;
  for (const RadialFn& f : rad)
    if (f.zc >= 0) m.radial_resolved = 1;
  m.nAG = (int)grp.size();
  m.maxmem = 1;
  for (const AngGroup& a : grp) m.maxmem = std::max(m.maxmem, a.nmem);
  m.all_lad = 1;
  for (AngGroup& a : grp) {
    ladder_form(a);
    a.partner = -1;
    a.follower = 0;
    if (a.lad == 0) m.all_lad = 0;
  }
  // G5 ladder groups that differ only in eta share one pass of the adjoint
  m.has_pairs = 0;
  for (size_t p = 0; p < grp.size(); ++p) {
    AngGroup& a = grp[p];
    if (a.type != 5 || a.lad == 0 || a.partner >= 0 || a.follower) continue;
    for (size_t q = p + 1; q < grp.size(); ++q) {
      AngGroup& b = grp[q];
      if (b.type != 5 || b.lad != a.lad || b.lmask != a.lmask || b.partner >= 0 || b.follower)
        continue;
      if (b.zc1 != a.zc1 || b.zc2 != a.zc2 || b.same != a.same || b.atomid != a.atomid ||
          b.rc != a.rc || b.lamv[0] != a.lamv[0] || b.lamv[1] != a.lamv[1])
        continue;
      a.partner = (int)q;
      b.follower = 1;
      m.has_pairs = 1;
      break;
    }
  }
  m.need_ext = 0;
  for (const RadialFn& f : rad) m.need_ext = std::max(m.need_ext, f.atomid);
  for (const AngGroup& a : grp) m.need_ext = std::max(m.need_ext, a.atomid);
  for (int k = 0; k < md->n_ext_in; ++k)
    if (md->ext_in_idx) m.need_ext = std::max(m.need_ext, md->ext_in_idx[k]);
  m.rcmax = rcmax;
  // network shape
  NetShape& sh = m.shape;
  sh.nl = md->n_layers;
  long long o = 0;
  for (int l = 0; l < sh.nl; ++l) sh.dims[l] = md->dims[l];
  for (int l = 0; l < sh.nl - 1; ++l) {
    sh.woff[l] = o;
    o += (long long)sh.dims[l] * sh.dims[l + 1];
  }
  sh.woff[sh.nl - 1] = o;
  o += sh.dims[sh.nl - 1];
  for (int l = 0; l < sh.nl; ++l) {
    sh.boff[l] = o;
    o += sh.dims[l];
  }
  sh.nparams = o;
  for (int sp = 0; sp < md->n_species; ++sp) {
    const int a = md->activation[sp];
    if (a < 0 || a > 9) return fail(ctx, FNET_ERR_ARG, "invalid activation id");
    sh.act[sp] = a;
  }
  std::vector<int> ext_in(md->n_ext_in);
  for (int k = 0; k < md->n_ext_in; ++k) {
    if (!md->ext_in_idx || md->ext_in_idx[k] < 1)
      return fail(ctx, FNET_ERR_ARG, "external input indices are 1-based");
    ext_in[k] = md->ext_in_idx[k] - 1;
  }
  const size_t np = (size_t)sh.nparams * md->n_species;
  m.h_params.assign(md->params, md->params + np);
  m.h_species_z.assign(md->species_z, md->species_z + md->n_species);
  hipStream_t s = ctx->stream;
  HIPCHECK(ctx, upload(&m.radial, rad.data(), rad.size(), s));
  HIPCHECK(ctx, upload(&m.groups, grp.data(), grp.size(), s));
  HIPCHECK(ctx, upload(&m.zmean, zmean.data(), zmean.size(), s));
  HIPCHECK(ctx, upload(&m.zinv, zinv.data(), zinv.size(), s));
  HIPCHECK(ctx, upload(&m.ext_in, ext_in.data(), ext_in.size(), s));
  HIPCHECK(ctx, upload(&m.params, m.h_params.data(), np, s));
  HIPCHECK(ctx, dalloc(&m.params_f, np));
  HIPCHECK(ctx, launch_to_float(m.params, m.params_f, (long long)np, s));
  HIPCHECK(ctx, hipStreamSynchronize(s));
  {
    const int prc = upload_wpack(ctx, m, s);
    if (prc) {
      free_model(m);
      return prc;
    }
  }
  m.valid = true;
  free_model(ctx->model);
  ctx->model = m;
  ctx->model_gen++;
  ctx->params_gen++;
  ctx->have_grad = false;
  return FNET_OK;
}

int fnet_params_upload(fnet_ctx* ctx, const double* params) {
  if (!ctx) return FNET_ERR_ARG;
  if (!ctx->model.valid) return fail(ctx, FNET_ERR_STATE, "no model uploaded");
  if (!params) return fail(ctx, FNET_ERR_ARG, "params is NULL");
  DevModel& m = ctx->model;
  const size_t np = (size_t)m.shape.nparams * m.nspecies;
  HIPCHECK(ctx, hipSetDevice(ctx->device));
  m.h_params.assign(params, params + np);
  HIPCHECK(ctx, hipMemcpyAsync(m.params, params, sizeof(double) * np, hipMemcpyHostToDevice,
                               ctx->stream));
  HIPCHECK(ctx, launch_to_float(m.params, m.params_f, (long long)np, ctx->stream));
  HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->params_gen++;
  return upload_wpack(ctx, m, ctx->stream);
}

}  // extern "C"

namespace {

// Lattice vectors of one periodic structure (column c = vector c): fnetdata.F90:1151-1154 rejects
// dependent vectors, and the image ranges of latpointiter.F90:226-253 must stay finite and modest
// (at most 1e6 lattice images), else the neighbour search would loop over them.
int check_lattice(fnet_ctx* ctx, const double* L, double rcmax, int s, long long natoms) {
  auto o = [&](int r, int c) { return L[c * 3 + r]; };
  for (int q = 0; q < 9; ++q)
    if (!std::isfinite(L[q]))
      return fail(ctx, FNET_ERR_ARG, "non-finite lattice vector in structure " +
                                         std::to_string(s));
  const double det = o(0, 0) * (o(1, 1) * o(2, 2) - o(2, 1) * o(1, 2)) -
                     o(0, 1) * (o(1, 0) * o(2, 2) - o(2, 0) * o(1, 2)) +
                     o(0, 2) * (o(1, 0) * o(2, 1) - o(2, 0) * o(1, 1));
  if (!(std::fabs(det) >= 1e-12))
    return fail(ctx, FNET_ERR_ARG, "dependent lattice vectors in structure " + std::to_string(s));
  double cof[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      cof[r][c] = o(r1, c1) * o(r2, c2) - o(r1, c2) * o(r2, c1);
    }
  double nimg = 1.0;
  for (int a = 0; a < 3; ++a) {
    // row a of inv(L) is column a of the cofactor matrix divided by det
    double ss = 0.0;
    for (int c = 0; c < 3; ++c) ss += (cof[c][a] / det) * (cof[c][a] / det);
    nimg *= 2.0 * (std::floor(rcmax * std::sqrt(ss)) + 1.0) + 1.0;
  }
  if (!(nimg <= 1.0e6))
    return fail(ctx, FNET_ERR_CAPACITY, "cell too small for the cutoff (more than 1e6 "
                                        "lattice images) in structure " + std::to_string(s));
  // neighbour entries carry the 32-bit key (image + 1) natoms + atom
  if (!(nimg * (double)natoms < 2147483647.0))
    return fail(ctx, FNET_ERR_CAPACITY, "lattice images x atoms exceed 2^31 in structure " +
                                            std::to_string(s));
  return FNET_OK;
}

// Validates a batch descriptor and loads it into b, reusing b's device buffers when they are
// large enough (fnet_batch_reload) or allocating them (fnet_batch_upload).
int load_batch(fnet_ctx* ctx, fnet_batch* b, const fnet_batch_desc* d) {
  if (!d) return fail(ctx, FNET_ERR_ARG, "NULL batch descriptor");
  if (!ctx->model.valid) return fail(ctx, FNET_ERR_STATE, "upload the model before a batch");
  if (d->n_struct < 0 || !d->offsets) return fail(ctx, FNET_ERR_ARG, "invalid structure count");
  const long long N = d->n_struct ? d->offsets[d->n_struct] : 0;
  if (N < 0 || N >= (1LL << 30)) return fail(ctx, FNET_ERR_ARG, "invalid atom count");
  if (d->offsets[0] != 0) return fail(ctx, FNET_ERR_ARG, "offsets[0] must be 0");
  for (int s = 0; s < d->n_struct; ++s)
    if (d->offsets[s + 1] < d->offsets[s]) return fail(ctx, FNET_ERR_ARG, "offsets decrease");
  if (N > 0 && (!d->coords || !d->species || !d->atnum))
    return fail(ctx, FNET_ERR_ARG, "coords/species/atnum missing");
  if (d->n_ext < 0 || (d->n_ext > 0 && !d->ext)) return fail(ctx, FNET_ERR_ARG, "invalid ext");
  const DevModel& m = ctx->model;
  if (m.need_ext > d->n_ext)
    return fail(ctx, FNET_ERR_ARG, "the model needs more external feature columns than given");
  HIPCHECK(ctx, hipSetDevice(ctx->device));
  const int S = d->n_struct;
  std::vector<int> per(S, 0), astruct(N), zc(N);
  std::vector<double> lat((size_t)9 * S, 0.0);
  for (int s = 0; s < S; ++s) {
    per[s] = d->periodic ? (d->periodic[s] != 0) : 0;
    if (per[s]) {
      if (!d->latvecs) return fail(ctx, FNET_ERR_ARG, "periodic structure without lattice vectors");
      for (int q = 0; q < 9; ++q) lat[9 * s + q] = d->latvecs[9 * s + q];
      const int lrc = check_lattice(ctx, &lat[9 * s], m.rcmax, s,
                                    d->offsets[s + 1] - d->offsets[s]);
      if (lrc) return lrc;
    }
    for (long long i = d->offsets[s]; i < d->offsets[s + 1]; ++i) astruct[i] = s;
  }
  int zone = m.nZ == 1;
  for (long long i = 0; i < N; ++i) {
    if (d->species[i] < 0 || d->species[i] >= m.nspecies)
      return fail(ctx, FNET_ERR_ARG, "species index out of range");
    zc[i] = zcode_of(m, d->atnum[i]);
    zone = zone && zc[i] == 0;
  }
  hipStream_t s = ctx->stream;
  HIPCHECK(ctx, hipStreamSynchronize(s));
  // capacity: atoms
  if (N > b->capN || !b->xyz) {
    void* pa[] = {b->xyz, b->atom_struct, b->species, b->zcode, b->cnt, b->rowptr, b->scan_tmp,
                  b->atom_cell, b->atom_m, b->cell_atoms};
    for (void* p : pa) dfree(p);
    const long long cap = std::max<long long>(N, 1);
    HIPCHECK(ctx, dalloc(&b->xyz, (size_t)3 * cap));
    HIPCHECK(ctx, dalloc(&b->atom_struct, (size_t)cap));
    HIPCHECK(ctx, dalloc(&b->species, (size_t)cap));
    HIPCHECK(ctx, dalloc(&b->zcode, (size_t)cap));
    HIPCHECK(ctx, dalloc(&b->cnt, (size_t)cap));
    HIPCHECK(ctx, dalloc(&b->rowptr, (size_t)cap + 1));
    HIPCHECK(ctx, dalloc(&b->scan_tmp, (size_t)scan_tmp_size(cap)));
    HIPCHECK(ctx, dalloc(&b->atom_cell, (size_t)cap));
    HIPCHECK(ctx, dalloc(&b->atom_m, (size_t)cap));
    HIPCHECK(ctx, dalloc(&b->cell_atoms, (size_t)cap));
    b->capN = cap;
  }
  if (N * d->n_ext > b->capExt || !b->ext) {
    dfree(b->ext);
    b->capExt = std::max<long long>(N * d->n_ext, 1);
    HIPCHECK(ctx, dalloc(&b->ext, (size_t)b->capExt));
  }
  // capacity: structures
  if (S > b->capS || !b->lat) {
    void* ps[] = {b->lat, b->periodic, b->off, b->sinfo, b->cellinfo, b->cell_off};
    for (void* p : ps) dfree(p);
    const long long cap = std::max(S, 1);
    HIPCHECK(ctx, dalloc(&b->lat, (size_t)9 * cap));
    HIPCHECK(ctx, dalloc(&b->periodic, (size_t)cap));
    HIPCHECK(ctx, dalloc(&b->off, (size_t)cap + 1));
    HIPCHECK(ctx, dalloc(&b->sinfo, (size_t)cap));
    HIPCHECK(ctx, dalloc(&b->cellinfo, (size_t)cap));
    HIPCHECK(ctx, dalloc(&b->cell_off, (size_t)cap + 1));
    b->capS = cap;
  }
  if (!b->d_maxnb) HIPCHECK(ctx, dalloc(&b->d_maxnb, 1));
  b->S = S;
  b->N = N;
  b->next = d->n_ext;
  b->h_off.assign(d->offsets, d->offsets + S + 1);
  b->h_species.assign(d->species, d->species + N);
  b->h_lat = lat;
  b->h_periodic = per;
  // chunks of the energy reduction: kEnergyChunk atoms of one structure each (the host tables
  // live until the stream synchronisation below: the copies may read them late)
  std::vector<int> cs, cf(S + 1, 0);
  {
    for (int st = 0; st < S; ++st) {
      cf[st] = (int)cs.size();
      const long long na = d->offsets[st + 1] - d->offsets[st];
      for (long long q = 0; q < na; q += kEnergyChunk) cs.push_back(st);
    }
    cf[S] = (int)cs.size();
    dfree(b->chunk_struct);
    dfree(b->chunk_first);
    b->chunk_struct = b->chunk_first = nullptr;
    HIPCHECK(ctx, upload(&b->chunk_struct, cs.data(), cs.size(), s));
    HIPCHECK(ctx, upload(&b->chunk_first, cf.data(), cf.size(), s));
    const long long need = std::max<long long>((long long)cs.size(), 1) * kMaxOut;
    if (need > b->epart_len) {
      dfree(b->Epart);
      b->Epart = nullptr;
      HIPCHECK(ctx, dalloc(&b->Epart, (size_t)need));
      b->epart_len = need;
    }
    b->nchunks = (int)cs.size();
    b->h_chunk_first = cf;
  }
  if (N) HIPCHECK(ctx, hipMemcpyAsync(b->xyz, d->coords, sizeof(double) * 3 * N,
                                      hipMemcpyHostToDevice, s));
  if (S) {
    HIPCHECK(ctx, hipMemcpyAsync(b->lat, lat.data(), sizeof(double) * lat.size(),
                                 hipMemcpyHostToDevice, s));
    HIPCHECK(ctx, hipMemcpyAsync(b->periodic, per.data(), sizeof(int) * S, hipMemcpyHostToDevice,
                                 s));
  }
  HIPCHECK(ctx, hipMemcpyAsync(b->off, d->offsets, sizeof(long long) * (S + 1),
                               hipMemcpyHostToDevice, s));
  if (N) {
    HIPCHECK(ctx, hipMemcpyAsync(b->atom_struct, astruct.data(), sizeof(int) * N,
                                 hipMemcpyHostToDevice, s));
    HIPCHECK(ctx, hipMemcpyAsync(b->species, d->species, sizeof(int) * N, hipMemcpyHostToDevice,
                                 s));
    HIPCHECK(ctx, hipMemcpyAsync(b->zcode, zc.data(), sizeof(int) * N, hipMemcpyHostToDevice, s));
  }
  b->h_atnum.assign(d->atnum, d->atnum + N);
  b->zone = zone;
  b->zcode_gen = ctx->model_gen;
  if (N * d->n_ext)
    HIPCHECK(ctx, hipMemcpyAsync(b->ext, d->ext, sizeof(double) * N * d->n_ext,
                                 hipMemcpyHostToDevice, s));
  HIPCHECK(ctx, hipStreamSynchronize(s));
  invalidate(b);
  b->tile_atoms = 0;  // species layout changed: rebuild the network tiles
  b->result_precision = -1;
  b->gen++;
  return setup_cells(b, lat.data(), per.data());
}

}  // namespace

extern "C" {

int fnet_batch_upload(fnet_ctx* ctx, const fnet_batch_desc* d, fnet_batch** out) {
  if (!ctx) return FNET_ERR_ARG;
  if (!d || !out) return fail(ctx, FNET_ERR_ARG, "NULL argument");
  *out = nullptr;
  fnet_batch* b = new fnet_batch();
  b->ctx = ctx;
  const int rc = load_batch(ctx, b, d);
  if (rc) {
    fnet_batch_destroy(b);
    return rc;
  }
  *out = b;
  return FNET_OK;
}

int fnet_batch_reload(fnet_batch* b, const fnet_batch_desc* d) {
  if (!b) return FNET_ERR_ARG;
  return load_batch(b->ctx, b, d);
}

int fnet_query(fnet_ctx* ctx) {
  if (!ctx) return FNET_ERR_ARG;
  const hipError_t e = hipStreamQuery(ctx->stream);
  if (e == hipSuccess) return 0;
  if (e == hipErrorNotReady) return 1;
  return hipfail(ctx, e, "hipStreamQuery");
}

void fnet_batch_destroy(fnet_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  void* ptrs[] = {b->xyz, b->lat, b->periodic, b->off, b->atom_struct, b->species, b->zcode,
                  b->ext, b->sinfo, b->perm, b->tile_sp, b->tile_start, b->tile_cnt, b->cnt,
                  b->rowptr, b->scan_tmp, b->nj, b->nimg, b->nd, b->d_maxnb, b->G, b->cG,
                  b->Eatom, b->Estruct, b->F, b->lossgrad, b->d_targets, b->cellinfo,
                  b->cell_off, b->cell_count, b->cell_start, b->cell_atoms, b->atom_cell,
                  b->atom_m, b->cell_scan_tmp, b->mtile_sp, b->mtile_start, b->mtile_cnt,
                  b->pnj, b->pnimg, b->pnd, b->gvec, b->mir, b->orph, b->d_norph,
                  b->chunk_struct, b->chunk_first, b->Epart, b->gpart};
  for (void* p : ptrs) dfree(p);
  drop_graph(b);
  if (b->h_maxnb) (void)hipHostFree(b->h_maxnb);
  if (b->ctx->last_train_batch == b) b->ctx->last_train_batch = nullptr;
  delete b;
}

int fnet_batch_update_coords(fnet_batch* b, const double* coords, const double* latvecs) {
  if (!b) return FNET_ERR_ARG;
  fnet_ctx* ctx = b->ctx;
  if (!coords) return fail(ctx, FNET_ERR_ARG, "coords is NULL");
  // the new cells pass the checks of a batch upload before anything is copied or launched
  if (latvecs)
    for (int st = 0; st < b->S; ++st)
      if (b->h_periodic[st]) {
        const int lrc = check_lattice(ctx, latvecs + 9 * st, ctx->model.rcmax, st,
                                      b->h_off[st + 1] - b->h_off[st]);
        if (lrc) return lrc;
      }
  HIPCHECK(ctx, hipSetDevice(ctx->device));
  HIPCHECK(ctx, hipMemcpyAsync(b->xyz, coords, sizeof(double) * 3 * b->N, hipMemcpyHostToDevice,
                               ctx->stream));
  if (latvecs) {
    HIPCHECK(ctx, hipMemcpyAsync(b->lat, latvecs, sizeof(double) * 9 * b->S,
                                 hipMemcpyHostToDevice, ctx->stream));
    for (int q = 0; q < 9 * b->S; ++q)
      if (b->h_periodic[q / 9]) b->h_lat[q] = latvecs[q];
    const int crc = setup_cells(b, b->h_lat.data(), b->h_periodic.data());
    if (crc) return crc;
  }
  HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
  invalidate(b);
  return FNET_OK;
}

int fnet_batch_set_mode(fnet_batch* b, int precision, int launch) {
  if (!b) return FNET_ERR_ARG;
  if (precision < FNET_FP64 || precision > FNET_BF16_FAST)
    return fail(b->ctx, FNET_ERR_ARG, "unknown precision");
  if (launch != FNET_BATCHED && launch != FNET_PER_STRUCT)
    return fail(b->ctx, FNET_ERR_ARG, "unknown launch mode");
  if (b->precision != precision || b->launch != launch) invalidate(b);
  b->precision = precision;
  b->launch = launch;
  return FNET_OK;
}

long long fnet_batch_n_atoms(const fnet_batch* b) { return b ? b->N : -1; }

int fnet_neighbour_list(fnet_ctx* ctx, fnet_batch* b, long long* n_pairs) {
  int rc = check_ready(ctx, b);
  if (rc) return rc;
  invalidate(b);
  rc = do_neighbours(b);
  if (rc) return rc;
  if (n_pairs) *n_pairs = b->npairs;
  return FNET_OK;
}

int fnet_get_neighbours(fnet_ctx* ctx, fnet_batch* b, long long* rowptr, int* j, int* img) {
  int rc = check_ready(ctx, b);
  if (rc) return rc;
  rc = validate_step(b);
  if (rc) return rc;
  if (!b->have_nl) return fail(ctx, FNET_ERR_STATE, "no neighbour list computed");
  if (b->nl_padded) {
    // row offsets of the padded rows (a steady-state step does not form them), then the rows
    // copied to them
    HIPCHECK(ctx, launch_scan(b->cnt, b->rowptr, b->N, b->scan_tmp, ctx->stream, nullptr));
    int P = 0;
    HIPCHECK(ctx, hipMemcpyAsync(&P, b->rowptr + b->N, sizeof(int), hipMemcpyDeviceToHost,
                                 ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (P > b->pcap) {
      dfree(b->nj);
      dfree(b->nimg);
      dfree(b->nd);
      b->nj = b->nimg = nullptr;
      b->nd = nullptr;
      const long long cap = (long long)P + P / 8 + 1024;
      HIPCHECK(ctx, dalloc(&b->nj, cap));
      HIPCHECK(ctx, dalloc(&b->nimg, cap));
      HIPCHECK(ctx, dalloc((double4**)&b->nd, cap));
      b->pcap = cap;
      b->gen++;
    }
    b->npairs = P;
    HIPCHECK(ctx, launch_compact_rows(b, b->result_precision == FNET_FP64, ctx->stream));
  }
  std::vector<int> rp(b->N + 1);
  HIPCHECK(ctx, hipMemcpyAsync(rp.data(), b->rowptr, sizeof(int) * (b->N + 1),
                               hipMemcpyDeviceToHost, ctx->stream));
  if (j && b->npairs)
    HIPCHECK(ctx, hipMemcpyAsync(j, b->nj, sizeof(int) * b->npairs, hipMemcpyDeviceToHost,
                                 ctx->stream));
  if (img && b->npairs)
    HIPCHECK(ctx, hipMemcpyAsync(img, b->nimg, sizeof(int) * b->npairs, hipMemcpyDeviceToHost,
                                 ctx->stream));
  HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (rowptr)
    for (long long i = 0; i <= b->N; ++i) rowptr[i] = rp[i];
  // the device list stores each entry's key (image + 1) n + (j - i) mod n: image = key / n - 1
  if (img)
    for (int st = 0; st < b->S; ++st) {
      const int n = (int)(b->h_off[st + 1] - b->h_off[st]);
      for (long long e = rp[b->h_off[st]]; e < rp[b->h_off[st + 1]]; ++e) img[e] = img[e] / n - 1;
    }
  return FNET_OK;
}

int fnet_acsf_forward(fnet_ctx* ctx, fnet_batch* b, double* G_out) {
  int rc = check_ready(ctx, b);
  if (rc) return rc;
  rc = validate_step(b);
  if (rc) return rc;
  if (!b->have_nl || b->result_precision != b->precision) {
    rc = do_neighbours(b);
    if (rc) return rc;
  }
  rc = do_acsf(b);
  if (rc) return rc;
  if (G_out)
    return download_T(ctx, b->G, 1, (size_t)b->N * ctx->model.nG, G_out);
  return FNET_OK;
}

int fnet_energy(fnet_ctx* ctx, fnet_batch* b, double* E_struct, double* E_atom) {
  int rc = check_ready(ctx, b);
  if (rc) return rc;
  rc = validate_step(b);
  if (rc) return rc;
  if (!b->have_G || b->result_precision != b->precision) {
    rc = fnet_acsf_forward(ctx, b, nullptr);
    if (rc) return rc;
  }
  if (!b->have_E) {
    rc = do_mlp(b, 0);
    if (rc) return rc;
    rc = do_reduce(b);
    if (rc) return rc;
  }
  const DevModel& m = ctx->model;
  if (E_struct)
    HIPCHECK(ctx, hipMemcpyAsync(E_struct, b->Estruct, sizeof(double) * b->S * m.nout,
                                 hipMemcpyDeviceToHost, ctx->stream));
  if (E_atom)
    HIPCHECK(ctx, hipMemcpyAsync(E_atom, b->Eatom, sizeof(double) * b->N * m.nout,
                                 hipMemcpyDeviceToHost, ctx->stream));
  HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
  return FNET_OK;
}

int fnet_forces(fnet_ctx* ctx, fnet_batch* b, double* F) {
  int rc = check_ready(ctx, b);
  if (rc) return rc;
  rc = validate_step(b);
  if (rc) return rc;
  if (!b->have_G || b->result_precision != b->precision) {
    rc = fnet_acsf_forward(ctx, b, nullptr);
    if (rc) return rc;
  }
  if (!b->have_F) {
    rc = do_mlp(b, 1);
    if (rc) return rc;
    rc = do_reduce(b);
    if (rc) return rc;
    rc = do_adjoint(b);
    if (rc) return rc;
  }
  if (F)
    return download_T(ctx, b->F, 1, (size_t)b->N * ctx->model.nout * 3, F);
  return FNET_OK;
}

int fnet_run(fnet_ctx* ctx, fnet_batch* b) {
  int rc = check_ready(ctx, b);
  if (rc) return rc;
  rc = ensure_tiles(b);
  if (rc) return rc;
  if (fast_step_ready(b)) return run_fast_step(b);
  return run_sync_step(b);
}

int fnet_device_memory(fnet_ctx* ctx, long long* free_bytes, long long* total_bytes) {
  if (!ctx || !free_bytes || !total_bytes) return FNET_ERR_ARG;
  HIPCHECK(ctx, hipSetDevice(ctx->device));
  size_t f = 0, t = 0;
  HIPCHECK(ctx, hipMemGetInfo(&f, &t));
  *free_bytes = (long long)f;
  *total_bytes = (long long)t;
  return FNET_OK;
}

int fnet_set_graph(fnet_ctx* ctx, int enable) {
  if (!ctx) return FNET_ERR_ARG;
  ctx->use_graph = enable ? 1 : 0;
  return FNET_OK;
}

int fnet_step_time(fnet_ctx* ctx, double* ms) {
  if (!ctx || !ms) return FNET_ERR_ARG;
  HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
  *ms = 0.0;
  if (ctx->n_ev < 2) return FNET_OK;
  float t = 0.f;
  HIPCHECK(ctx, hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[ctx->n_ev - 1]));
  *ms = t;
  return FNET_OK;
}

int fnet_batch_step_info(const fnet_batch* b, int* graph) {
  if (!b || !graph) return FNET_ERR_ARG;
  *graph = b->last_graph;
  return FNET_OK;
}

int fnet_stage_times(fnet_ctx* ctx, double* ms, int n) {
  if (!ctx || !ms) return FNET_ERR_ARG;
  HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
  int k = 0;
  for (; k < n && k + 1 < ctx->n_ev; ++k) {
    float t = 0.f;
    HIPCHECK(ctx, hipEventElapsedTime(&t, ctx->ev[k], ctx->ev[k + 1]));
    ms[k] = t;
  }
  return FNET_OK;
}

namespace {
bool tune_valid(const int* k) {
  return (k[0] == 0 || k[0] == 1 || k[0] == 2 || k[0] == 4) && (k[1] >= 0 && k[1] <= 2) &&
         (k[2] == 0 || k[2] == 1 || k[2] == 2 || k[2] == 4) && (k[3] >= -1 && k[3] <= 1) &&
         (k[4] == 0 || k[4] == 1);
}
}  // namespace

int fnet_set_tune(fnet_ctx* ctx, fnet_batch* b, const int* knobs) {
  if (!ctx || !b || !knobs) return FNET_ERR_ARG;
  if (b->ctx != ctx) return fail(ctx, FNET_ERR_ARG, "batch belongs to another context");
  if (b->launch != FNET_BATCHED) return fail(ctx, FNET_ERR_ARG, "tuning applies to batched launches");
  if (!tune_valid(knobs)) return fail(ctx, FNET_ERR_ARG, "invalid launch parameters");
  Tune t;
  t.fwd_wpb = knobs[0];
  t.adj_wpb = knobs[1];
  t.mlp_waves = knobs[2];
  t.rbal = knobs[3];
  t.persist = knobs[4];
  ctx->tunes[tune_key(b)] = t;
  return FNET_OK;
}

int fnet_get_tune(fnet_ctx* ctx, const fnet_batch* b, int* knobs) {
  if (!ctx || !b || !knobs) return FNET_ERR_ARG;
  const Tune& t = tune_of(b);
  knobs[0] = t.fwd_wpb;
  knobs[1] = t.adj_wpb;
  knobs[2] = t.mlp_waves;
  knobs[3] = t.rbal;
  knobs[4] = t.persist;
  return FNET_OK;
}

int fnet_autotune(fnet_ctx* ctx, fnet_batch* b, int reps, double* ms, int* knobs) {
  int rc = check_ready(ctx, b);
  if (rc) return rc;
  if (b->launch != FNET_BATCHED) return fail(ctx, FNET_ERR_ARG, "tuning applies to batched launches");
  reps = std::max(1, std::min(reps, 50));
  const int saved_graph = ctx->use_graph;
  ctx->use_graph = 0;  // eager steps, with the stage events
  const std::vector<long long> key = tune_key(b);
  ctx->tunes.erase(key);
  // time between events k0 and k1 of the steady-state step: the least of reps steps
  auto measure = [&](int k0, int k1, double* out) -> int {
    double best = 1e300;
    for (int r = 0; r < reps; ++r) {
      int e = fnet_run(ctx, b);
      if (e) return e;
      HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
      const bool eager = !b->last_graph && ctx->n_ev == 6;
      e = validate_step(b);
      if (e) return e;
      if (!eager) continue;
      float t = 0.f;
      HIPCHECK(ctx, hipEventElapsedTime(&t, ctx->ev[k0], ctx->ev[k1]));
      best = std::min(best, (double)t);
    }
    *out = best;
    return FNET_OK;
  };
  double t_default = 0.0, t_tuned = 0.0, dummy = 0.0;
  // two warm-up steps (the first sizes the list capacities), then the defaults
  rc = measure(0, 5, &dummy);
  if (!rc) rc = measure(0, 5, &dummy);
  if (!rc) rc = measure(0, 5, &t_default);
  Tune t;
  // coordinate descent over the knobs, each on the stages it changes
  auto choose = [&](int Tune::*field, std::initializer_list<int> cands, int k0, int k1) -> int {
    int best_v = t.*field;
    double best_t = 1e300;
    for (int v : cands) {
      t.*field = v;
      ctx->tunes[key] = t;
      double x = 0.0;
      const int e = measure(k0, k1, &x);
      if (e) return e;
      if (x < best_t) {
        best_t = x;
        best_v = v;
      }
    }
    t.*field = best_v;
    ctx->tunes[key] = t;
    return FNET_OK;
  };
  if (!rc) rc = choose(&Tune::fwd_wpb, {2, 1, 4}, 1, 2);
  if (!rc) rc = choose(&Tune::rbal, {ctx->model.radial_resolved, 1 - ctx->model.radial_resolved}, 1, 2);
  if (!rc) rc = choose(&Tune::adj_wpb, {1, 2}, 3, 4);
  if (!rc) rc = choose(&Tune::persist, {0, 1}, 1, 4);
  if (!rc && ctx->model.wpack && !net64(b->precision) && mlp_mfma_waves(ctx->model.shape))
    rc = choose(&Tune::mlp_waves, {mlp_mfma_waves(ctx->model.shape), 2, 1}, 2, 3);
  if (!rc) rc = measure(0, 5, &t_tuned);
  // per-stage choices that do not add up to a faster step (noise, interactions) fall back to
  // the defaults
  if (!rc && t_tuned >= t_default) {
    ctx->tunes.erase(key);
    rc = measure(0, 5, &t_tuned);
  }
  ctx->use_graph = saved_graph;
  if (rc) {
    ctx->tunes.erase(key);
    return rc;
  }
  if (ms) {
    ms[0] = t_default;
    ms[1] = t_tuned;
  }
  return fnet_get_tune(ctx, b, knobs);
}

int fnet_train_gradients(fnet_ctx* ctx, fnet_batch* b, int n_global, const double* g_targets,
                         const double* a_targets, const double* atomic_weights,
                         const double* dp_weights, double* grad, double* pred,
                         double* loss) {
  int rc = check_ready(ctx, b);
  if (rc) return rc;
  rc = validate_step(b);
  if (rc) return rc;
  const DevModel& m = ctx->model;
  if (n_global < 0 || n_global > m.nout) return fail(ctx, FNET_ERR_ARG, "invalid n_global");
  if (n_global > 0 && !g_targets) return fail(ctx, FNET_ERR_ARG, "g_targets missing");
  if (n_global < m.nout && !a_targets) return fail(ctx, FNET_ERR_ARG, "a_targets missing");
  if (!b->have_G || b->result_precision != b->precision) {
    rc = fnet_acsf_forward(ctx, b, nullptr);
    if (rc) return rc;
  }
  rc = do_mlp(b, 0);
  if (rc) return rc;
  rc = do_reduce(b);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  // stage inputs: gt [S*ng], at [N*na], aw [N], dpw [S]; then the loss, a pad, the
  // per-structure loss terms [2 S] and the per-chunk partial sums [3 nchunks]
  const long long na = m.nout - n_global;
  const long long need =
      (long long)b->S * n_global + b->N * na + b->N + 3LL * b->S + 3LL * b->nchunks + 4;
  if (need > b->targets_cap) {
    dfree(b->d_targets);
    b->d_targets = nullptr;
    HIPCHECK(ctx, dalloc(&b->d_targets, (size_t)need));
    b->targets_cap = need;
  }
  double* gt = b->d_targets;
  double* at = gt + (long long)b->S * n_global;
  double* aw = at + b->N * na;
  double* dpw = aw + b->N;
  double wsum = 0.0;
  if (n_global > 0)
    HIPCHECK(ctx, hipMemcpyAsync(gt, g_targets, sizeof(double) * b->S * n_global,
                                 hipMemcpyHostToDevice, s));
  if (na > 0)
    HIPCHECK(ctx, hipMemcpyAsync(at, a_targets, sizeof(double) * b->N * na, hipMemcpyHostToDevice,
                                 s));
  if (atomic_weights)
    HIPCHECK(ctx, hipMemcpyAsync(aw, atomic_weights, sizeof(double) * b->N,
                                 hipMemcpyHostToDevice, s));
  if (dp_weights) {
    HIPCHECK(ctx, hipMemcpyAsync(dpw, dp_weights, sizeof(double) * b->S, hipMemcpyHostToDevice, s));
    for (int k = 0; k < b->S; ++k) wsum += dp_weights[k];
  } else {
    wsum = (double)b->S;
  }
  if (!(wsum > 0.0)) return fail(ctx, FNET_ERR_ARG, "datapoint weights sum to zero");
  HIPCHECK(ctx, launch_lossgrad(b, m, n_global, gt, at, atomic_weights ? aw : nullptr,
                                dp_weights ? dpw : nullptr, 1.0 / wsum, ctx->loss, s));
  const long long np = m.shape.nparams * m.nspecies;
  if (np > ctx->grad_cap) {
    dfree(ctx->d_grad);
    ctx->d_grad = nullptr;
    HIPCHECK(ctx, dalloc(&ctx->d_grad, (size_t)np + m.nspecies));
    ctx->grad_cap = np;
  }
  {
    // per-tile partials, then the per-chunk sums of every species
    const int parts = grad_parts(b, m, net64(b->precision));
    const long long need = std::max<long long>((long long)parts, 1) * m.shape.nparams +
                           (long long)grad_chunks(parts) * m.shape.nparams * m.nspecies;
    if (need > b->gpart_len) {
      dfree(b->gpart);
      b->gpart = nullptr;
      HIPCHECK(ctx, dalloc(&b->gpart, (size_t)need));
      b->gpart_len = need;
    }
  }
  HIPCHECK(ctx, launch_mlp_grad(b, m, net64(b->precision), ctx->d_grad, s));
  if (ctx->reg_strength > 0.0)
    HIPCHECK(ctx, launch_regularize(m, ctx->d_grad, ctx->reg_strength / wsum, ctx->reg_alpha, s));
  double* dloss = dpw + b->S;
  HIPCHECK(ctx, launch_loss(b, m, n_global, gt, at, atomic_weights ? aw : nullptr,
                            dp_weights ? dpw : nullptr, ctx->loss, dloss, dloss + 2,
                            dloss + 2 + 2LL * b->S, s));
  if (grad)
    HIPCHECK(ctx, hipMemcpyAsync(grad, ctx->d_grad, sizeof(double) * np, hipMemcpyDeviceToHost, s));
  if (pred)
    HIPCHECK(ctx, hipMemcpyAsync(pred, b->Eatom, sizeof(double) * b->N * m.nout,
                                 hipMemcpyDeviceToHost, s));
  double lv = 0.0;
  HIPCHECK(ctx, hipMemcpyAsync(&lv, dloss, sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHECK(ctx, hipStreamSynchronize(s));
  if (loss) *loss = lv;
  ctx->last_loss = lv;
  ctx->have_grad = true;
  ctx->last_train_batch = b;
  return FNET_OK;
}

int fnet_set_loss(fnet_ctx* ctx, int loss) {
  if (!ctx) return FNET_ERR_ARG;
  if (loss < FNET_LOSS_MSE || loss > FNET_LOSS_RMS) return fail(ctx, FNET_ERR_ARG, "invalid loss");
  ctx->loss = loss;
  return FNET_OK;
}

int fnet_set_regularization(fnet_ctx* ctx, double strength, double alpha) {
  if (!ctx) return FNET_ERR_ARG;
  if (!(strength >= 0.0)) return fail(ctx, FNET_ERR_ARG, "regularisation strength must be >= 0");
  if (!(alpha >= 0.0 && alpha <= 1.0))
    return fail(ctx, FNET_ERR_ARG, "elastic net alpha out of bounds [0, 1]");
  ctx->reg_strength = strength;
  ctx->reg_alpha = strength > 0.0 ? alpha : 0.0;
  return FNET_OK;
}

int fnet_optimizer_step_sd(fnet_ctx* ctx, double lr, double max_disp, double tol,
                           double* params_out, int* converged) {
  if (!ctx) return FNET_ERR_ARG;
  if (!ctx->model.valid) return fail(ctx, FNET_ERR_STATE, "no model uploaded");
  if (!ctx->have_grad) return fail(ctx, FNET_ERR_STATE, "no gradient computed");
  if (!(max_disp > 0.0)) return fail(ctx, FNET_ERR_ARG, "max_disp must be positive");
  DevModel& m = ctx->model;
  HIPCHECK(ctx, hipSetDevice(ctx->device));
  const long long np = m.shape.nparams * m.nspecies;
  double* gmax = ctx->d_grad + np;
  HIPCHECK(ctx, launch_sd_step(m.params, ctx->d_grad, m.shape.nparams, lr, max_disp, tol, gmax,
                               m.nspecies, ctx->stream));
  ctx->params_gen++;
  HIPCHECK(ctx, launch_to_float(m.params, m.params_f, np, ctx->stream));
  std::vector<double> gm(m.nspecies);
  HIPCHECK(ctx, hipMemcpyAsync(m.h_params.data(), m.params, sizeof(double) * np,
                               hipMemcpyDeviceToHost, ctx->stream));
  HIPCHECK(ctx, hipMemcpyAsync(gm.data(), gmax, sizeof(double) * m.nspecies,
                               hipMemcpyDeviceToHost, ctx->stream));
  HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
  {
    const int prc = upload_wpack(ctx, m, ctx->stream);
    if (prc) return prc;
  }
  if (params_out) std::memcpy(params_out, m.h_params.data(), sizeof(double) * np);
  if (converged) {
    int c = 1;
    for (double v : gm) c = c && (v < tol);
    *converged = c;
  }
  ctx->have_grad = false;
  return FNET_OK;
}

int fnet_optimizer_init(fnet_ctx* ctx, int kind, double tol, double min_disp, double max_disp,
                        double learning_rate, int memory, int line_search) {
  if (!ctx) return FNET_ERR_ARG;
  if (!ctx->model.valid) return fail(ctx, FNET_ERR_STATE, "no model uploaded");
  if (kind < FNET_OPT_SD || kind > FNET_OPT_FIRE) return fail(ctx, FNET_ERR_ARG, "unknown optimizer");
  if (!(max_disp > 0.0)) return fail(ctx, FNET_ERR_ARG, "max_disp must be positive");
  (void)min_disp;  // the reference's MinDisplacement does not enter these updates
  const int mem = kind == FNET_OPT_LBFGS ? std::max(1, memory) : 1;
  DevModel& m = ctx->model;
  HIPCHECK(ctx, hipSetDevice(ctx->device));
  const long long n = m.shape.nparams;
  const long long vstride = opt_vec_doubles(n, mem);
  const long long need = vstride * m.nspecies;
  if (need > ctx->opt_cap || !ctx->opt_states) {
    dfree(ctx->opt_vecs);
    dfree(ctx->opt_states);
    dfree(ctx->opt_conv);
    ctx->opt_vecs = nullptr;
    ctx->opt_states = nullptr;
    ctx->opt_conv = nullptr;
    HIPCHECK(ctx, dalloc(&ctx->opt_vecs, (size_t)need));
    HIPCHECK(ctx, dalloc(&ctx->opt_states, 64));
    HIPCHECK(ctx, dalloc(&ctx->opt_conv, 64));
    ctx->opt_cap = need;
  }
  OptState st{};
  st.kind = kind;
  st.converged = 0;
  st.mem = mem;
  st.line_search = line_search ? 1 : 0;
  st.tol = tol;
  st.max_disp = max_disp;
  st.lr = learning_rate;
  // the line minimiser of CG takes 10 iterations and tolerance 1e4, that of L-BFGS 10 and tol
  st.lm_maxiter = 10;
  st.lm_tol = kind == FNET_OPT_CG ? 10000.0 : tol;
  st.cg_state = 1;
  st.lb_it = 0;
  st.f_it = 0;
  st.f_reset = 0;
  st.f_a = 0.1;                 // a_start
  st.f_dt = 0.1 * max_disp;     // dt_init = 0.1 dt_max
  std::vector<OptState> h(m.nspecies, st);
  hipStream_t s = ctx->stream;
  HIPCHECK(ctx, hipMemcpyAsync(ctx->opt_states, h.data(), sizeof(OptState) * h.size(),
                               hipMemcpyHostToDevice, s));
  HIPCHECK(ctx, launch_opt_init(m.params, ctx->opt_vecs, n, vstride, mem, m.nspecies, s));
  HIPCHECK(ctx, hipStreamSynchronize(s));
  ctx->opt_kind = kind;
  ctx->opt_vstride = vstride;
  return FNET_OK;
}

int fnet_optimizer_step(fnet_ctx* ctx, int kind, double* params_out, int* converged) {
  if (!ctx) return FNET_ERR_ARG;
  if (!ctx->model.valid) return fail(ctx, FNET_ERR_STATE, "no model uploaded");
  if (ctx->opt_kind < 0) return fail(ctx, FNET_ERR_STATE, "optimizer not initialised");
  if (kind != ctx->opt_kind)
    return fail(ctx, FNET_ERR_ARG, "optimizer kind differs from fnet_optimizer_init");
  if (!ctx->have_grad) return fail(ctx, FNET_ERR_STATE, "no gradient computed");
  DevModel& m = ctx->model;
  HIPCHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const long long n = m.shape.nparams, np = n * m.nspecies;
  HIPCHECK(ctx, launch_opt_step(ctx->opt_states, m.params, ctx->d_grad, ctx->opt_vecs, n,
                                ctx->opt_vstride, ctx->last_loss, ctx->opt_conv, m.nspecies, s));
  ctx->params_gen++;
  HIPCHECK(ctx, launch_to_float(m.params, m.params_f, np, s));
  std::vector<int> conv(m.nspecies, 0);
  HIPCHECK(ctx, hipMemcpyAsync(m.h_params.data(), m.params, sizeof(double) * np,
                               hipMemcpyDeviceToHost, s));
  HIPCHECK(ctx, hipMemcpyAsync(conv.data(), ctx->opt_conv, sizeof(int) * m.nspecies,
                               hipMemcpyDeviceToHost, s));
  HIPCHECK(ctx, hipStreamSynchronize(s));
  {
    const int prc = upload_wpack(ctx, m, s);
    if (prc) return prc;
  }
  if (params_out) std::memcpy(params_out, m.h_params.data(), sizeof(double) * np);
  if (converged) *converged = conv[m.nspecies - 1];
  ctx->have_grad = false;
  return FNET_OK;
}

int fnet_sync(fnet_ctx* ctx) {
  if (!ctx) return FNET_ERR_ARG;
  HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
  return FNET_OK;
}

int fnet_batch_sync(fnet_batch* b) {
  if (!b) return FNET_ERR_ARG;
  HIPCHECK(b->ctx, hipSetDevice(b->ctx->device));
  return validate_step(b);
}

}  // extern "C"
