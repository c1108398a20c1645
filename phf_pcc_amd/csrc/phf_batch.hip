// Batched projected energy + gradient over many independent instances.
//
// Each instance (Hamiltonian in factored form, spin grid, N electrons) owns a
// context that shares the batch's stream.  One evaluation of the batch runs,
// on that stream:
//   for every instance  C0 upload, K1/K2, density pack and split     (eval_pre)
//   one ERI GEMM launch over all instances: the 16-bit forms go through a
//       reduced forms (FP32 included) go through a persistent grouped kernel
//       over a work list of (instance, row tile, column tile, contraction
//       part) items; FP64 launches per instance
//   for every instance  K4/K5 grid sums (eval_post) and the J = 0 finish:
//       E0 = Re H / |Re S| (EigenSolver of the 1 x 1 problem, PHFTools.py:8-18)
//       and the (NSO-N) x N gradient block (Ysum - E0 Xsum)[N:, :] / S
//       (LocalG with fsp = fpg = fk = 1, PHFTools.f90:203-247)
//   results to pinned host memory.
// The first evaluation of a precision runs eagerly (allocating workspaces and
// autotuning the grouped tiles); the same launch sequence is then captured in
// a hipGraph and replayed by later evaluations, with only the pinned C0 block
// rewritten on the host.  The sequence records no memset nodes: zero fills
// are kernels.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <numeric>

#include "phf_internal.h"

namespace {

// out[0..7] = E0, Re H, Im H, Re S, Im S, status, 0, 0; out[8 ..] = the
// gradient block, C order (NSO-N) x N complex.  status: 0 ok, 1 + g for the
// first singular grid point g, -1 for |Re S| <= 1e-6 (no projected state).
__global__ __launch_bounds__(256) void batch_finish_kernel(int nso, int N, int G, const int* __restrict__ info,
                                                           const double* __restrict__ hs,
                                                           const double* __restrict__ ysum,
                                                           const double* __restrict__ xsum, double* __restrict__ out) {
  const double hr = hs[0], hi = hs[1], sr = hs[2], si = hs[3];
  const bool ok = fabs(sr) > 1e-6;
  const double e0 = ok ? hr / fabs(sr) : 0.0;
  if (blockIdx.x == 0) {
    __shared__ int first;
    if (threadIdx.x == 0) first = G;
    __syncthreads();
    for (int g = threadIdx.x; g < G; g += blockDim.x)
      if (info[g] != 0) atomicMin(&first, g);
    __syncthreads();
    if (threadIdx.x == 0) {
      out[0] = e0;
      out[1] = hr;
      out[2] = hi;
      out[3] = sr;
      out[4] = si;
      out[5] = first < G ? 1.0 + first : (ok ? 0.0 : -1.0);
      out[6] = 0.0;
      out[7] = 0.0;
    }
  }
  const int nel = (nso - N) * N;
  const double den = sr * sr + si * si;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nel; t += gridDim.x * blockDim.x) {
    const int a = N + t / N, i = t % N;
    const long q = 2 * ((long)a + (long)i * nso);
    const double vr = ysum[q] - e0 * xsum[q], vi = ysum[q + 1] - e0 * xsum[q + 1];
    // v / S
    out[8 + 2 * t] = ok ? (vr * sr + vi * si) / den : 0.0;
    out[8 + 2 * t + 1] = ok ? (vi * sr - vr * si) / den : 0.0;
  }
}

}  // namespace

struct phf_batch_s {
  int device = 0;
  int ncu = 256;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<phf_ctx*> inst;
  std::vector<int> nocc;
  // pinned host staging and the device result block
  double* h_c0 = nullptr;
  double* h_out = nullptr;
  double* d_out = nullptr;
  size_t n_c0 = 0, n_out = 0;
  std::vector<size_t> off_c0, off_out;
  // plan of the current precision
  int plan_prec = -1;
  std::vector<EvalBufs> bufs;
  // grouped 16-bit ERI GEMM
  H16Item* d_items = nullptr;
  size_t items_cap = 0;
  int nitems = 0, cfg = 0, splitk = 1, grid = 0;
  std::string tune_log;
  int autotune = 1;
  // graph of one evaluation
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // side streams (default priority) that run the instances' per-point stages
  // concurrently; fork/join events order them around the grouped GEMM
  static constexpr int kSide = 8;
  hipStream_t side[kSide] = {};
  hipEvent_t fork_ev[2] = {};
  hipEvent_t join_ev[2][kSide] = {};
  // timing of the last evaluation
  hipEvent_t ev[4] = {};
  double ms_total = 0, ms_gemm = 0;
  bool last_graph = false;
};

namespace {

int bfail(phf_batch* bt, int code, const std::string& msg) {
  if (bt) bt->err = msg;
  return code;
}

#define BHIPCK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return bfail(bt, e_ == hipErrorOutOfMemory ? PHF_ERR_NOMEM : PHF_ERR_HIP,       \
                   std::string(#expr) + ": " + hipGetErrorString(e_));                \
  } while (0)

void drop_graph(phf_batch* bt) {
  if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
  if (bt->graph) (void)hipGraphDestroy(bt->graph);
  bt->exec = nullptr;
  bt->graph = nullptr;
}

// Work list of the grouped launch over every instance (shared builder).
void build_items(phf_batch* bt, int prec, int cfg, int sk, int grid, std::vector<H16Item>& out) {
  std::vector<std::pair<const phf_ctx*, const EvalBufs*>> units;
  for (size_t i = 0; i < bt->inst.size(); ++i) units.push_back({bt->inst[i], &bt->bufs[i]});
  build_h16_items(units, prec, cfg, sk, grid, out);
}

int upload_items(phf_batch* bt, const std::vector<H16Item>& items) {
  if (items.size() > bt->items_cap) {
    if (bt->d_items) BHIPCK(hipFree(bt->d_items));
    bt->d_items = nullptr;
    BHIPCK(hipMalloc(&bt->d_items, items.size() * sizeof(H16Item)));
    bt->items_cap = items.size();
  }
  BHIPCK(hipMemcpy(bt->d_items, items.data(), items.size() * sizeof(H16Item), hipMemcpyHostToDevice));
  bt->nitems = (int)items.size();
  return PHF_OK;
}

// Zero fills (split contractions) and the grouped launch of the chosen work list.
void launch_grouped(phf_batch* bt, int prec) {
  if (bt->splitk > 1)
    for (size_t i = 0; i < bt->inst.size(); ++i) {
      const EvalBufs& b = bt->bufs[i];
      if (b.rjk)
        fill_zero(bt->stream, b.Cjk, (b.Mp + b.Mk) * b.kP);
      else
        fill_zero(bt->stream, b.Gam, b.Mp * bt->inst[i]->nrpad);
    }
  eri_gemm_mixed_grouped(bt->stream, prec, bt->d_items, bt->nitems, bt->cfg, bt->grid, bt->splitk > 1);
}

// Choose (cfg, split-K) of the grouped launch on the live operands (the
// density splits of every instance must be in place): each candidate's work
// list runs once to warm up and twice timed with HIP events, the fastest wins.
int tune_grouped(phf_batch* bt, int prec) {
  std::vector<std::pair<int, int>> cands;
  const int forced = h16_forced_config();
  for (int c = 0; c < h16_num_configs(); ++c) {
    if (forced >= 0 ? c != forced : !h16_cfg_fold64(c)) continue;
    bool fits = true;   // A rows (dense or per block) are multiples of 256
    for (phf_ctx* x : bt->inst) {
      fits = fits && (x->nrpad % h16_cfg_bn(c) == 0) && !(h16_cfg_rn16(c) && x->nso > h16_rn16_max_nso());
      for (int q = 0; q < 3; ++q)   // every contraction length of the instance (dense, each block, spin-free)
        fits = fits && h16_cfg_serves(c, prec, phf_spinfree(x) ? x->rj_kj
                                                               : (x->blocked && x->use_blocks) ? x->blk_k[q] : x->kt);
    }
    if (!fits) continue;
    for (int sk : {1, 2})
      if (!h16_forced_splitk() || sk == h16_forced_splitk()) cands.push_back({c, sk});
  }
  if (cands.empty()) return bfail(bt, PHF_ERR_UNSUPPORTED, "phf_batch: no grouped tile config fits");
  std::pair<int, int> best = cands.front();
  std::vector<H16Item> items;
  char line[200];
  std::string log = "{\"grouped\": true, \"prec\": " + std::to_string(prec) + ", \"candidates\": [";
  if (bt->autotune && cands.size() > 1) {
    float best_ms = 1e30f;
    for (size_t q = 0; q < cands.size(); ++q) {
      bt->cfg = cands[q].first;
      bt->splitk = cands[q].second;
      bt->grid = h16_grouped_grid(bt->cfg, bt->ncu);
      build_items(bt, prec, bt->cfg, bt->splitk, bt->grid, items);
      bt->grid = std::min<int>(bt->grid, (int)items.size());
      {
        int rc = upload_items(bt, items);
        if (rc) return rc;
      }
      launch_grouped(bt, prec);
      float ms_min = 1e30f;
      for (int rep = 0; rep < 2; ++rep) {
        BHIPCK(hipEventRecord(bt->ev[0], bt->stream));
        launch_grouped(bt, prec);
        BHIPCK(hipEventRecord(bt->ev[1], bt->stream));
        BHIPCK(hipEventSynchronize(bt->ev[1]));
        float ms;
        BHIPCK(hipEventElapsedTime(&ms, bt->ev[0], bt->ev[1]));
        ms_min = std::min(ms_min, ms);
      }
      snprintf(line, sizeof line, "%s{\"cfg\": %d, \"splitk\": %d, \"items\": %zu, \"ms\": %.4f}", q ? ", " : "",
               cands[q].first, cands[q].second, items.size(), ms_min);
      log += line;
      if (ms_min < best_ms) {
        best_ms = ms_min;
        best = cands[q];
      }
    }
  }
  bt->cfg = best.first;
  bt->splitk = best.second;
  bt->grid = h16_grouped_grid(bt->cfg, bt->ncu);
  build_items(bt, prec, bt->cfg, bt->splitk, bt->grid, items);
  bt->grid = std::min<int>(bt->grid, (int)items.size());
  snprintf(line, sizeof line, "], \"chosen\": {\"cfg\": %d, \"splitk\": %d, \"grid\": %d, \"items\": %zu}}\n",
           bt->cfg, bt->splitk, bt->grid, items.size());
  bt->tune_log += log + line;
  return upload_items(bt, items);
}

// Instances run their per-point stages on the side streams: instance i (in
// order of decreasing work) on side stream i % kSide, with its context's
// stream pointed there for the duration of the enqueue.
static void fork(phf_batch* bt, int phase) {
  (void)hipEventRecord(bt->fork_ev[phase], bt->stream);
  for (int k = 0; k < phf_batch_s::kSide; ++k) (void)hipStreamWaitEvent(bt->side[k], bt->fork_ev[phase], 0);
}

static void join(phf_batch* bt, int phase) {
  for (int k = 0; k < phf_batch_s::kSide; ++k) {
    (void)hipEventRecord(bt->join_ev[phase][k], bt->side[k]);
    (void)hipStreamWaitEvent(bt->stream, bt->join_ev[phase][k], 0);
  }
}

// The launch sequence of one evaluation after the plan (no allocation, no
// host synchronisation: capturable).  Events only when `timed`.
int enqueue_eval(phf_batch* bt, int prec, bool timed) {
  hipStream_t s = bt->stream;
  const size_t ni = bt->inst.size();
  std::vector<int> order(ni);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return bt->bufs[a].G * bt->inst[a]->nso > bt->bufs[b].G * bt->inst[b]->nso;
  });
  auto on_side = [&](size_t rank, phf_ctx* c) { c->stream = bt->side[rank % phf_batch_s::kSide]; };
  if (timed) BHIPCK(hipEventRecord(bt->ev[0], s));
  fork(bt, 0);
  for (size_t r = 0; r < ni; ++r) {
    const int i = order[r];
    phf_ctx* c = bt->inst[i];
    const EvalBufs& b = bt->bufs[i];
    on_side(r, c);
    BHIPCK(hipMemcpyAsync(b.C0, bt->h_c0 + bt->off_c0[i], sizeof(double) * 2 * (size_t)c->nso * b.N,
                          hipMemcpyHostToDevice, c->stream));
    eval_pre(c, b);
    if (prec == PHF_PREC_FP64) {
      int rc = eval_gemm(c, b);   // per-instance GEMMs overlap on the side streams
      if (rc) {
        for (phf_ctx* x : bt->inst) x->stream = bt->stream;
        return bfail(bt, rc, c->err);
      }
    }
  }
  join(bt, 0);
  if (timed) BHIPCK(hipEventRecord(bt->ev[1], s));
  if (prec != PHF_PREC_FP64) launch_grouped(bt, prec);
  if (timed) BHIPCK(hipEventRecord(bt->ev[2], s));
  fork(bt, 1);
  for (size_t r = 0; r < ni; ++r) {
    const int i = order[r];
    phf_ctx* c = bt->inst[i];
    const EvalBufs& b = bt->bufs[i];
    on_side(r, c);
    if (prec != PHF_PREC_FP64 && b.rjk)   // Gamma + h from the grouped products (split-K 1:
      rjk_unpack(c, b, c->stream, bt->splitk > 1);   // Gamma_ab, Gamma_ba were written directly)
    eval_post(c, b, true);
    const int nel = (c->nso - b.N) * b.N;
    const int blocks = std::max(1, std::min(1024, (nel + 255) / 256));
    batch_finish_kernel<<<blocks, 256, 0, c->stream>>>(c->nso, b.N, (int)b.G, b.Info, b.HS, b.Ysum, b.Xsum,
                                                       bt->d_out + bt->off_out[i]);
  }
  join(bt, 1);
  for (phf_ctx* c : bt->inst) c->stream = bt->stream;
  BHIPCK(hipGetLastError());
  BHIPCK(hipMemcpyAsync(bt->h_out, bt->d_out, sizeof(double) * bt->n_out, hipMemcpyDeviceToHost, s));
  if (timed) BHIPCK(hipEventRecord(bt->ev[3], s));
  return PHF_OK;
}

}  // namespace

extern "C" int phf_batch_create(int device, int ninst, phf_batch** out) {
  if (!out) return PHF_ERR_ARG;
  *out = nullptr;
  if (ninst <= 0 || ninst > 4096) return PHF_ERR_ARG;
  phf_batch* bt = new phf_batch_s();
  phf_ctx probe;
  int rc = phf_ctx_init(&probe, device, nullptr);
  if (rc) {
    delete bt;
    return rc;
  }
  bt->device = device;
  bt->ncu = probe.ncu;
  bt->stream = probe.stream;   // the batch owns the stream created for the probe
  bt->inst.assign(ninst, nullptr);
  bt->nocc.assign(ninst, 0);
  bool ok = true;
  for (auto& e : bt->ev) ok = ok && hipEventCreateWithFlags(&e, kTimingEventFlags) == hipSuccess;
  for (auto& e : bt->fork_ev) ok = ok && hipEventCreateWithFlags(&e, kSyncEventFlags) == hipSuccess;
  for (auto& row : bt->join_ev)
    for (auto& e : row) ok = ok && hipEventCreateWithFlags(&e, kSyncEventFlags) == hipSuccess;
  for (auto& st : bt->side) ok = ok && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
  if (!ok) {
    (void)phf_batch_destroy(bt);
    return PHF_ERR_HIP;
  }
  *out = bt;
  return PHF_OK;
}

extern "C" int phf_batch_destroy(phf_batch* bt) {
  if (!bt) return PHF_OK;
  (void)hipSetDevice(bt->device);
  (void)hipStreamSynchronize(bt->stream);
  drop_graph(bt);
  for (phf_ctx* c : bt->inst)
    if (c) (void)phf_ctx_destroy(c);
  if (bt->d_items) (void)hipFree(bt->d_items);
  if (bt->d_out) (void)hipFree(bt->d_out);
  if (bt->h_c0) (void)hipHostFree(bt->h_c0);
  if (bt->h_out) (void)hipHostFree(bt->h_out);
  for (auto& e : bt->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : bt->fork_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& row : bt->join_ev)
    for (auto& e : row)
      if (e) (void)hipEventDestroy(e);
  for (auto& st : bt->side)
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  if (bt->stream) (void)hipStreamDestroy(bt->stream);
  delete bt;
  return PHF_OK;
}

extern "C" const char* phf_batch_last_error(const phf_batch* bt) { return bt ? bt->err.c_str() : ""; }

extern "C" int phf_batch_set_instance(phf_batch* bt, int i, const double* h, const double* b, int n, int nl,
                                      const double* ref, int nocc, const int* ngrid, const double* r1, int nr1,
                                      const double* r2, int nr2, const double* rpg, int npg,
                                      const double* weightsp, const double* weightpg, int storage, int flags) {
  if (!bt) return PHF_ERR_ARG;
  if (i < 0 || i >= (int)bt->inst.size()) return bfail(bt, PHF_ERR_ARG, "phf_batch_set_instance: index out of range");
  if (nocc <= 0 || nocc >= 2 * n) return bfail(bt, PHF_ERR_ARG, "phf_batch_set_instance: nocc out of range");
  if (storage < PHF_PREC_FP64 || storage > PHF_PREC_I8X4)
    return bfail(bt, PHF_ERR_ARG, "phf_batch_set_instance: unknown storage precision");
  BHIPCK(hipSetDevice(bt->device));
  BHIPCK(hipStreamSynchronize(bt->stream));
  drop_graph(bt);
  bt->plan_prec = -1;
  if (bt->inst[i]) {
    (void)phf_ctx_destroy(bt->inst[i]);
    bt->inst[i] = nullptr;
  }
  phf_ctx* c = new phf_ctx_s();
  int rc = phf_ctx_init(c, bt->device, bt->stream);
  auto done = [&](int code) -> int {
    if (code) {
      bt->err = "instance " + std::to_string(i) + ": " + c->err;
      (void)phf_ctx_destroy(c);
      return code;
    }
    bt->inst[i] = c;
    bt->nocc[i] = nocc;
    return PHF_OK;
  };
  if (rc) return done(rc);
  rc = phf_set_hamiltonian_factored(c, h, b, n, nl, ref, (flags & PHF_HAM_DENSE) ? PHF_HAM_DENSE : PHF_HAM_BLOCKS_ONLY);
  if (rc) return done(rc);
  // J = 0 spin grid: nci = ncisp = ncipg = 1, no Wigner factors
  const double one = 1.0;
  std::vector<double> wpg(npg > 0 ? npg : 1, 1.0);
  rc = phf_set_grid(c, ref, ngrid, 1, 1, r1, nr1, r2, nr2, rpg, npg, weightsp, weightpg ? weightpg : wpg.data(), 1,
                    &one, 1, &one, 1, &one, 1, 0, 1);
  if (rc) return done(rc);
  rc = storage != PHF_PREC_FP64 ? phf_keep_only_mixed(c, storage) : phf_keep_used_structure(c);
  if (rc) return done(rc);
  rc = phf_release_ws(c);
  return done(rc);
}

extern "C" int phf_batch_set_graph(phf_batch* bt, int on) {
  if (!bt) return PHF_ERR_ARG;
  bt->use_graph = on != 0;
  if (!bt->use_graph) drop_graph(bt);
  return PHF_OK;
}

extern "C" int phf_batch_autotune(phf_batch* bt, int on) {
  if (!bt) return PHF_ERR_ARG;
  bt->autotune = on != 0;
  drop_graph(bt);
  bt->plan_prec = -1;
  return PHF_OK;
}

// Plan every instance for `prec`, size the staging blocks, run one eager
// evaluation (autotuning the grouped GEMM), then capture the sequence.
static int plan_batch(phf_batch* bt, int prec) {
  drop_graph(bt);
  bt->plan_prec = -1;
  const size_t ni = bt->inst.size();
  bt->bufs.assign(ni, EvalBufs());
  bt->off_c0.assign(ni, 0);
  bt->off_out.assign(ni, 0);
  size_t nc0 = 0, nout = 0;
  for (size_t i = 0; i < ni; ++i) {
    phf_ctx* c = bt->inst[i];
    int rc = eval_plan(c, bt->nocc[i], prec, 1, bt->bufs[i]);
    if (rc) return bfail(bt, rc, "instance " + std::to_string(i) + ": " + c->err);
    bt->off_c0[i] = nc0;
    nc0 += 2 * (size_t)c->nso * bt->nocc[i];
    nout = (nout + 31) / 32 * 32;   // 256-byte aligned result blocks
    bt->off_out[i] = nout;
    nout += 8 + 2 * (size_t)(c->nso - bt->nocc[i]) * bt->nocc[i];
  }
  if (nc0 > bt->n_c0) {
    if (bt->h_c0) BHIPCK(hipHostFree(bt->h_c0));
    bt->h_c0 = nullptr;
    BHIPCK(hipHostMalloc(&bt->h_c0, sizeof(double) * nc0, hipHostMallocDefault));
    bt->n_c0 = nc0;
  }
  if (nout > bt->n_out) {
    if (bt->h_out) BHIPCK(hipHostFree(bt->h_out));
    if (bt->d_out) BHIPCK(hipFree(bt->d_out));
    bt->h_out = nullptr;
    bt->d_out = nullptr;
    BHIPCK(hipHostMalloc(&bt->h_out, sizeof(double) * nout, hipHostMallocDefault));
    BHIPCK(hipMalloc(&bt->d_out, sizeof(double) * nout));
    bt->n_out = nout;
  }
  bt->plan_prec = prec;
  return PHF_OK;
}

extern "C" int phf_batch_eval(phf_batch* bt, const double* z, int cmplxconj, int precision, double* energies,
                              double* grads, double* hs, int* status) {
  if (!bt) return PHF_ERR_ARG;
  if (!z || !energies) return bfail(bt, PHF_ERR_ARG, "phf_batch_eval: NULL argument");
  if (cmplxconj != 0 && cmplxconj != 1)
    return bfail(bt, PHF_ERR_UNSUPPORTED, "phf_batch_eval: CmplxConj must be 0 or 1");
  if (precision < PHF_PREC_FP64 || precision > PHF_PREC_I8X4)
    return bfail(bt, PHF_ERR_ARG, "phf_batch_eval: unknown precision");
  for (size_t i = 0; i < bt->inst.size(); ++i) {
    if (!bt->inst[i]) return bfail(bt, PHF_ERR_ARG, "phf_batch_eval: instance " + std::to_string(i) + " not set");
    const int a = bt->inst[i]->a_only;
    if (a && a != precision)
      return bfail(bt, PHF_ERR_UNSUPPORTED,
                   "phf_batch_eval: instance " + std::to_string(i) + " stores A for another precision only");
  }
  BHIPCK(hipSetDevice(bt->device));
  const bool fresh = bt->plan_prec != precision;
  if (fresh) {
    int rc = plan_batch(bt, precision);
    if (rc) return rc;
  }
  // C0 blocks of every instance into the pinned staging block
  size_t zoff = 0;
  for (size_t i = 0; i < bt->inst.size(); ++i) {
    const int nso = bt->inst[i]->nso, N = bt->nocc[i];
    build_c0(z + zoff, nso, N, cmplxconj, bt->h_c0 + bt->off_c0[i]);
    zoff += 2 * (size_t)(nso - N) * N;
  }
  hipStream_t s = bt->stream;
  const bool h16 = precision != PHF_PREC_FP64;   // every reduced form runs one grouped launch
  if (fresh || !bt->exec) {
    if (fresh && h16) {
      // the grouped tiles are tuned on live operands: densities first
      for (size_t i = 0; i < bt->inst.size(); ++i) {
        BHIPCK(hipMemcpyAsync(bt->bufs[i].C0, bt->h_c0 + bt->off_c0[i],
                              sizeof(double) * 2 * (size_t)bt->inst[i]->nso * bt->nocc[i], hipMemcpyHostToDevice, s));
        eval_pre(bt->inst[i], bt->bufs[i]);
      }
      int rc = tune_grouped(bt, precision);
      if (rc) return rc;
    }
    int rc = enqueue_eval(bt, precision, true);
    if (rc) return rc;
    BHIPCK(hipStreamSynchronize(s));
    float t03, t12;
    BHIPCK(hipEventElapsedTime(&t03, bt->ev[0], bt->ev[3]));
    BHIPCK(hipEventElapsedTime(&t12, bt->ev[1], bt->ev[2]));
    bt->ms_total = t03;
    bt->ms_gemm = t12;
    bt->last_graph = false;
    if (bt->use_graph) {
      BHIPCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      rc = enqueue_eval(bt, precision, false);
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(s, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      BHIPCK(ec);
      bt->graph = g;
      BHIPCK(hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0));
    }
  } else {
    BHIPCK(hipEventRecord(bt->ev[0], s));
    BHIPCK(hipGraphLaunch(bt->exec, s));
    BHIPCK(hipEventRecord(bt->ev[3], s));
    BHIPCK(hipStreamSynchronize(s));
    float t03;
    BHIPCK(hipEventElapsedTime(&t03, bt->ev[0], bt->ev[3]));
    bt->ms_total = t03;
    bt->last_graph = true;
  }
  // unpack; singular points, empty projection and the Hermiticity test
  // (PHFTools.f90:195-200 at nci = 1: |2 Im H| against 1e-5 |H|) fail an
  // instance; every instance's outputs are written, the first failure is
  // returned and status[i] (if given) holds each instance's code
  int first_rc = PHF_OK;
  size_t goff = 0;
  for (size_t i = 0; i < bt->inst.size(); ++i) {
    const phf_ctx* c = bt->inst[i];
    const double* o = bt->h_out + bt->off_out[i];
    const int N = bt->nocc[i];
    const double st = o[5];
    int rc = PHF_OK;
    std::string msg;
    if (st > 0.5) {
      const long g = (long)st - 1;
      const long ny = c->g_ny, npg = c->g_npg;
      char buf[200];
      snprintf(buf, sizeof buf,
               "Error in ZGetRF: instance %zu: singular overlap at grid point (il=%ld, iy=%ld, ipg=%ld)", i,
               g / (ny * npg) + 1, (g / npg) % ny + 1, g % npg + 1);
      rc = PHF_ERR_SINGULAR;
      msg = buf;
    } else if (st < -0.5) {
      rc = PHF_ERR_SINGULAR;
      msg = "instance " + std::to_string(i) + ": projected overlap below 1e-6";
    } else if (std::fabs(2 * o[2]) > 1e-5 * std::hypot(o[1], o[2]) ||
               std::fabs(2 * o[4]) > 1e-5 * std::hypot(o[3], o[4])) {
      rc = PHF_ERR_NOT_HERMITIAN;
      msg = "instance " + std::to_string(i) + ": Not Hermitian";
    }
    if (rc && !first_rc) {
      first_rc = rc;
      bt->err = msg;
    }
    if (status) status[i] = rc;
    energies[i] = o[0];
    if (hs) std::memcpy(hs + 4 * i, o + 1, 4 * sizeof(double));
    const size_t ng = 2 * (size_t)(c->nso - N) * N;
    if (grads) std::memcpy(grads + goff, o + 8, ng * sizeof(double));
    goff += ng;
  }
  return first_rc;
}

extern "C" int phf_batch_stats(const phf_batch* bt, double* ms, int nmax) {
  if (!bt || !ms || nmax <= 0) return 0;
  int nsf = 0;
  for (const phf_ctx* c : bt->inst) nsf += c ? phf_spinfree(c) : 0;
  const double v[4] = {bt->ms_total, bt->last_graph ? -1.0 : bt->ms_gemm, bt->last_graph ? 1.0 : 0.0, (double)nsf};
  const int n = std::min(nmax, 4);
  for (int i = 0; i < n; ++i) ms[i] = v[i];
  return n;
}

extern "C" int phf_batch_tuning_report(const phf_batch* bt, char* buf, int len) {
  if (!bt || !buf || len <= 0) return 0;
  const int n = std::min<int>(len - 1, (int)bt->tune_log.size());
  std::memcpy(buf, bt->tune_log.data(), n);
  buf[n] = 0;
  return n;
}
