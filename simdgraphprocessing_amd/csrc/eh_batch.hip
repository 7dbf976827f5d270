// eh_batch.hip -- many independent graphs in one device arena.
//
// Graph g of a batch gets n_g vertices and m_g Erdos-Renyi draws (eh_batch_params)
// from the stream key (seed, 1 + g); its vertices occupy ids [vbase_g, vbase_g + n_g)
// of one union graph, relabelled by degree inside that range only.  Every edge
// stays inside its graph, so one oriented matrix over the union answers all
// graphs at once and each row carries its graph id for the per-graph counters.
//
// Counting modes:
//   persistent  triangle + 4-clique passes over the whole arena (the edge-centric
//               kernel for short rows and the persistent task kernels) captured once
//               into a hipGraph and replayed as one launch
//   per_graph   the same kernels launched separately for every graph's rows
#include <algorithm>
#include <numeric>

#include "eh_internal.hpp"
#include "eh_rng.hpp"

namespace {

inline unsigned grid_for(uint64_t work, unsigned block = 256) {
  uint64_t g = ceil_div(work ? work : 1, block);
  return (unsigned)(g > 65536 ? 65536 : g);
}

__global__ void k_gen_batch(uint64_t *keys, uint64_t total, const uint64_t *moff, const uint32_t *nv,
                            const uint32_t *vbase, uint32_t ng, uint64_t seed) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = ng - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (moff[mid] <= e) lo = mid; else hi = mid - 1;
    }
    const uint32_t g = lo;
    const uint64_t r = eh_draw(eh_stream_key(seed, 1ull + g), e - moff[g]);
    const uint32_t n = nv[g], b = vbase[g];
    const uint32_t u = eh_scale32((uint32_t)r, n), v = eh_scale32((uint32_t)(r >> 32), n);
    uint64_t k = 0xFFFFFFFFFFFFFFFFull;
    if (u != v) {
      const uint32_t x = b + (u < v ? u : v), y = b + (u < v ? v : u);
      k = ((uint64_t)x << 32) | y;
    }
    keys[e] = k;
  }
}

__global__ void k_vertex_graph(const uint32_t *vbase, uint32_t ng, uint32_t nt, uint32_t *vg) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nt; v += gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = ng - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (vbase[mid] <= v) lo = mid; else hi = mid - 1;
    }
    vg[v] = lo;
  }
}

__global__ void k_zero64(unsigned long long *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

// the error words of the triangle and 4-clique passes (queue-head blocks 0..7 and 8..15,
// word 7 of each; a fill kernel, so no memset node is recorded into the captured graph)
__global__ void k_zero_err(uint32_t *nx) {
  if (threadIdx.x == 0) {
    nx[7] = 0u;
    nx[15] = 0u;
  }
}

}  // namespace

// what a captured graph or a per-graph plan bakes in: the schedule buffers it reads
// and the launch shape it was recorded with
struct PassKey {
  bool want = false;
  int sched = -1;
  TuneParams tp;
  const void *tasks = nullptr, *small_rows = nullptr;
  uint64_t n_big = 0, n_small = 0, n_small_rows = 0;
  uint32_t lds_small = 0, lds_big = 0;
};

static bool same_tp(const TuneParams &x, const TuneParams &y) {
  return x.lanes == y.lanes && x.task_work == y.task_work && x.lds_small == y.lds_small && x.lds_big == y.lds_big &&
         x.waves_per_block == y.waves_per_block && x.big_waves == y.big_waves && x.small_deg == y.small_deg &&
         x.stream_ratio == y.stream_ratio && x.bstat_j == y.bstat_j && x.grab == y.grab &&
         x.local_max == y.local_max && x.bsh_max == y.bsh_max && x.short_queues == y.short_queues;
}

static bool same_key(const PassKey &x, const PassKey &y) {
  if (x.want != y.want) return false;
  if (!x.want) return true;
  return x.sched == y.sched && same_tp(x.tp, y.tp) && x.tasks == y.tasks && x.small_rows == y.small_rows &&
         x.n_big == y.n_big && x.n_small == y.n_small && x.n_small_rows == y.n_small_rows &&
         x.lds_small == y.lds_small && x.lds_big == y.lds_big;
}

struct eh_batch {
  eh_ctx *ctx = nullptr;
  uint32_t ng = 0;
  uint32_t nv_total = 0;
  std::vector<uint32_t> n, vbase;
  std::vector<uint64_t> m;
  eh_graph *g = nullptr;
  eh_matrix *M = nullptr;
  DevBuf vgraph, counts, next;
  // per-graph launch plans (tasks grouped by graph, per class)
  struct Plan {
    bool valid = false;
    PassKey key;
    DevBuf small_rows, tasks_big, tasks_small, local_rows[2];
    std::vector<uint64_t> sr_off, tb_off, ts_off, lr_off[2];  // ng + 1 offsets
  } plan[2];  // [0] triangles, [1] 4-cliques
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  PassKey graph_key[2];
  ~eh_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    eh_matrix_free(M);
    eh_graph_free(g);
  }
};

extern "C" eh_status eh_batch_generate(eh_ctx *ctx, uint64_t seed, uint32_t ngraphs, eh_layout_policy policy,
                                       eh_batch **out) {
  if (!ctx || !out || ngraphs == 0) { eh_set_error("eh_batch_generate: bad argument"); return EH_ERR_INVALID; }
  if (ngraphs >= (1u << 12)) {
    if (ngraphs > (1u << 12)) { eh_set_error("at most 4096 graphs per batch"); return EH_ERR_UNSUPPORTED; }
  }
  EH_HIP(hipSetDevice(ctx->device));
  eh_batch *b = new eh_batch();
  b->ctx = ctx;
  b->ng = ngraphs;
  b->n.resize(ngraphs);
  b->m.resize(ngraphs);
  b->vbase.resize(ngraphs);
  std::vector<uint64_t> moff(ngraphs + 1, 0);
  uint64_t vt = 0;
  for (uint32_t g = 0; g < ngraphs; g++) {
    eh_batch_params(seed, g, &b->n[g], &b->m[g]);
    b->vbase[g] = (uint32_t)vt;
    vt += b->n[g];
    moff[g + 1] = moff[g] + b->m[g];
  }
  if (vt >= (1ull << 26)) { delete b; eh_set_error("batch too large for the grouped relabel"); return EH_ERR_UNSUPPORTED; }
  b->nv_total = (uint32_t)vt;
  hipStream_t st = ctx->stream;
  DevBuf dmoff, dn, dvb, keys;
  eh_status s;
  auto fail = [&](eh_status e) { delete b; return e; };
  if ((s = dmoff.alloc(8ull * (ngraphs + 1))) != EH_OK || (s = dn.alloc(4ull * ngraphs)) != EH_OK ||
      (s = dvb.alloc(4ull * ngraphs)) != EH_OK || (s = keys.alloc(8ull * moff[ngraphs] + 8)) != EH_OK ||
      (s = b->vgraph.alloc(4ull * vt + 4)) != EH_OK)
    return fail(s);
  if (hipMemcpy(dmoff.p, moff.data(), 8ull * (ngraphs + 1), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dn.p, b->n.data(), 4ull * ngraphs, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dvb.p, b->vbase.data(), 4ull * ngraphs, hipMemcpyHostToDevice) != hipSuccess)
    return fail(EH_ERR_HIP);
  k_gen_batch<<<grid_for(moff[ngraphs]), 256, 0, st>>>(keys.as<uint64_t>(), moff[ngraphs], dmoff.as<uint64_t>(),
                                                       dn.as<uint32_t>(), dvb.as<uint32_t>(), ngraphs, seed);
  k_vertex_graph<<<grid_for(vt), 256, 0, st>>>(dvb.as<uint32_t>(), ngraphs, (uint32_t)vt, b->vgraph.as<uint32_t>());
  if (hipGetLastError() != hipSuccess) return fail(EH_ERR_HIP);
  if ((s = graph_canonicalize(ctx, (uint32_t)vt, keys, moff[ngraphs], 1, &b->g, b->vgraph.as<uint32_t>())) != EH_OK)
    return fail(s);
  if ((s = matrix_build_dev(ctx, b->g->n, b->g->off.as<uint64_t>(), b->g->nbr.as<uint32_t>(), b->g->entries,
                            EH_FILTER_LOWER, nullptr, policy, b->vgraph.as<uint32_t>(), b->g->symmetric,
                            &b->M)) != EH_OK)
    return fail(s);
  if ((s = b->counts.alloc(16ull * ngraphs + 64)) != EH_OK || (s = b->next.alloc(64)) != EH_OK) return fail(s);
  *out = b;
  return EH_OK;
}

extern "C" eh_status eh_batch_rebuild(eh_batch *b, eh_layout_policy policy, const uint8_t *row_types) {
  if (!b || !b->g) { eh_set_error("eh_batch_rebuild: null batch"); return EH_ERR_INVALID; }
  eh_ctx *ctx = b->ctx;
  EH_HIP(hipSetDevice(ctx->device));
  DevBuf dt;
  if (row_types) {
    // per-row arms as eh_matrix_build_typed takes them (the LEARNED policy's guard applies)
    EH_CHECK(dt.alloc((size_t)b->nv_total + 16));
    EH_HIP(hipMemcpy(dt.p, row_types, b->nv_total, hipMemcpyHostToDevice));
    policy = EH_LAYOUT_LEARNED;
  }
  eh_matrix *M = nullptr;
  EH_CHECK(matrix_build_dev(ctx, b->g->n, b->g->off.as<uint64_t>(), b->g->nbr.as<uint32_t>(), b->g->entries,
                            EH_FILTER_LOWER, nullptr, policy, b->vgraph.as<uint32_t>(), b->g->symmetric, &M,
                            row_types ? dt.as<uint8_t>() : nullptr));
  // the captured graph and the per-graph plans read the old matrix's buffers: record them again
  if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  b->graph_key[0] = b->graph_key[1] = PassKey();
  b->plan[0].valid = b->plan[1].valid = false;
  eh_matrix_free(b->M);
  b->M = M;
  return EH_OK;
}

extern "C" eh_status eh_batch_matrix(eh_batch *b, eh_matrix **out) {
  if (!b || !out || !b->M) { eh_set_error("eh_batch_matrix: null argument"); return EH_ERR_INVALID; }
  *out = b->M;
  return EH_OK;
}

extern "C" eh_status eh_batch_info(eh_batch *b, uint32_t *ngraphs, uint64_t *total_vertices,
                                   uint64_t *total_edges) {
  if (!b) { eh_set_error("null batch"); return EH_ERR_INVALID; }
  if (ngraphs) *ngraphs = b->ng;
  if (total_vertices) *total_vertices = b->nv_total;
  if (total_edges) *total_edges = b->g->entries / 2;
  return EH_OK;
}

extern "C" eh_status eh_batch_graph_params(eh_batch *b, uint32_t *n, uint64_t *m_draws) {
  if (!b) { eh_set_error("null batch"); return EH_ERR_INVALID; }
  for (uint32_t g = 0; g < b->ng; g++) {
    if (n) n[g] = b->n[g];
    if (m_draws) m_draws[g] = b->m[g];
  }
  return EH_OK;
}

extern "C" eh_status eh_batch_graph_download(eh_batch *b, uint32_t g, uint64_t *row_off, uint32_t *nbrs,
                                             uint64_t cap) {
  if (!b || !row_off || g >= b->ng) { eh_set_error("bad argument"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(b->ctx->device));
  const uint32_t n = b->n[g], v0 = b->vbase[g];
  std::vector<uint64_t> off(n + 1);
  EH_HIP(hipMemcpy(off.data(), b->g->off.as<uint64_t>() + v0, 8ull * (n + 1), hipMemcpyDeviceToHost));
  const uint64_t e = off[n] - off[0];
  if (e > cap || (e && !nbrs)) { eh_set_error("graph %u needs %llu entries", g, (unsigned long long)e); return EH_ERR_CAPACITY; }
  if (e) EH_HIP(hipMemcpy(nbrs, b->g->nbr.as<uint32_t>() + off[0], 4 * e, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i <= n; i++) row_off[i] = off[i] - off[0];
  for (uint64_t k = 0; k < e; k++) nbrs[k] -= v0;
  return EH_OK;
}

// group a union schedule's work by graph for the per-graph launches
static eh_status build_plan(eh_batch *b, int kidx, const PassKey &key) {
  eh_batch::Plan &P = b->plan[kidx];
  if (P.valid && same_key(P.key, key)) return EH_OK;
  P.valid = false;
  eh_matrix *M = b->M;
  Schedule &s = M->sched[kidx][key.sched];
  const uint32_t ng = b->ng;
  std::vector<uint32_t> sr(s.n_small_rows);
  std::vector<uint4> tasks(s.n_big + s.n_small);
  if (!sr.empty()) EH_HIP(hipMemcpy(sr.data(), s.small_rows.p, 4 * sr.size(), hipMemcpyDeviceToHost));
  if (!tasks.empty()) EH_HIP(hipMemcpy(tasks.data(), s.tasks.p, 16 * tasks.size(), hipMemcpyDeviceToHost));
  auto gid_row = [&](uint32_t r) {
    return (uint32_t)(std::upper_bound(b->vbase.begin(), b->vbase.end(), r) - b->vbase.begin() - 1);
  };
  auto group = [&](std::vector<uint4> v, std::vector<uint64_t> &off) {
    std::stable_sort(v.begin(), v.end(), [](const uint4 &x, const uint4 &y) {
      return (x.w & 0x0FFFFFFFu) < (y.w & 0x0FFFFFFFu);
    });
    off.assign(ng + 1, 0);
    for (const uint4 &t : v) off[(t.w & 0x0FFFFFFFu) + 1]++;
    for (uint32_t g = 0; g < ng; g++) off[g + 1] += off[g];
    return v;
  };
  std::vector<uint4> tb(tasks.begin(), tasks.begin() + s.n_big), ts(tasks.begin() + s.n_big, tasks.end());
  tb = group(tb, P.tb_off);
  ts = group(ts, P.ts_off);
  P.sr_off.assign(ng + 1, 0);
  for (uint32_t r : sr) P.sr_off[gid_row(r) + 1]++;
  for (uint32_t g = 0; g < ng; g++) P.sr_off[g + 1] += P.sr_off[g];
  // the local-graph row classes (4-cliques), in row order and so already grouped by graph
  for (int c = 0; c < 2; c++) {
    std::vector<uint32_t> lr(s.n_local[c]);
    if (!lr.empty()) EH_HIP(hipMemcpy(lr.data(), s.local_rows[c].p, 4 * lr.size(), hipMemcpyDeviceToHost));
    P.lr_off[c].assign(ng + 1, 0);
    for (uint32_t r : lr) P.lr_off[c][gid_row(r) + 1]++;
    for (uint32_t g = 0; g < ng; g++) P.lr_off[c][g + 1] += P.lr_off[c][g];
    EH_CHECK(P.local_rows[c].alloc(4 * lr.size() + 16));
    if (!lr.empty()) EH_HIP(hipMemcpy(P.local_rows[c].p, lr.data(), 4 * lr.size(), hipMemcpyHostToDevice));
  }
  EH_CHECK(P.small_rows.alloc(4 * sr.size() + 16));
  EH_CHECK(P.tasks_big.alloc(16 * tb.size() + 16));
  EH_CHECK(P.tasks_small.alloc(16 * ts.size() + 16));
  if (!sr.empty()) EH_HIP(hipMemcpy(P.small_rows.p, sr.data(), 4 * sr.size(), hipMemcpyHostToDevice));
  if (!tb.empty()) EH_HIP(hipMemcpy(P.tasks_big.p, tb.data(), 16 * tb.size(), hipMemcpyHostToDevice));
  if (!ts.empty()) EH_HIP(hipMemcpy(P.tasks_small.p, ts.data(), 16 * ts.size(), hipMemcpyHostToDevice));
  P.valid = true;
  P.key = key;
  return EH_OK;
}

extern "C" eh_status eh_batch_count(eh_batch *b, int mode, const eh_run_opts *tri_opts,
                                    const eh_run_opts *k4_opts, uint64_t *tri, uint64_t *k4, double *ms) {
  if (!b || (mode != 0 && mode != 1)) { eh_set_error("eh_batch_count: bad argument"); return EH_ERR_INVALID; }
  eh_ctx *ctx = b->ctx;
  EH_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  eh_matrix *M = b->M;
  const uint32_t ng = b->ng;
  const eh_run_opts *opts[2] = {tri_opts, k4_opts};
  const bool want[2] = {tri != nullptr, k4 != nullptr};
  const int kk[2] = {3, 4};
  PassKey key[2];
  for (int q = 0; q < 2; q++) {
    key[q].want = want[q];
    if (!want[q]) continue;
    const int sched = opts[q] ? opts[q]->schedule : EH_SCHED_TUNED;
    if (sched != EH_SCHED_PLAIN && sched != EH_SCHED_TUNED) { eh_set_error("bad schedule %d", sched); return EH_ERR_INVALID; }
    // the pass's schedule, built or reused (no pass runs here: the counts come from the graph
    // or the per-graph launches below, whose error words are checked after them)
    EH_CHECK(schedule_ensure(M, kk[q], opts[q], nullptr));
    const Schedule &S = M->sched[q][sched];
    key[q].sched = sched;
    key[q].tp = tune_defaults(M, sched, kk[q], opts[q]);
    key[q].tasks = S.tasks.p;
    key[q].small_rows = S.small_rows.p;
    key[q].n_big = S.n_big;
    key[q].n_small = S.n_small;
    key[q].n_small_rows = S.n_small_rows;
    key[q].lds_small = S.lds_small;
    key[q].lds_big = S.lds_big;
  }
  // Both passes in one: when both are tuned and every row of both is a short row (no tasks, no
  // local-graph rows, no B-stationary part), the two short-row lists hold the same rows (one
  // rule, one bound), and the 4-clique kernel's local-edge hits are the triangles of those rows:
  // k_small_cliques4 then counts both and the triangle pass is not launched
  bool fuse = false;
  if (want[0] && want[1] && key[0].sched == EH_SCHED_TUNED && key[1].sched == EH_SCHED_TUNED) {
    const Schedule &s3 = M->sched[0][EH_SCHED_TUNED], &s4 = M->sched[1][EH_SCHED_TUNED];
    fuse = s3.n_big == 0 && s3.n_small == 0 && s3.n_btasks == 0 && s3.n_htasks == 0 && s3.bstat_j == 0 &&
           s4.n_big == 0 && s4.n_small == 0 && s4.n_local[0] == 0 && s4.n_local[1] == 0 &&
           s3.small_deg == s4.small_deg && s3.n_small_rows == s4.n_small_rows && s3.n_small_rows > 0;
  }
  unsigned long long *cnt = b->counts.as<unsigned long long>();
  uint32_t *nx = b->next.as<uint32_t>();
  float elapsed = 0;
  if (mode == 0) {
    // one captured graph: zero the counters, then the triangle and 4-clique passes; it is
    // recorded again whenever a pass's schedule buffers or launch shape changed
    if (!b->exec || !same_key(b->graph_key[0], key[0]) || !same_key(b->graph_key[1], key[1])) {
      if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
      if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
      EH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      k_zero64<<<grid_for(2ull * ng), 256, 0, st>>>(cnt, 2 * ng);
      k_zero_err<<<1, 64, 0, st>>>(nx);
      eh_status s = EH_OK;
      for (int q = 0; q < 2 && s == EH_OK; q++)
        if (want[q] && !(fuse && q == 0))
          s = schedule_launch(M, kk[q], key[q].sched, key[q].tp, M->sched[q][key[q].sched],
                              b->counts.as<uint64_t>() + (uint64_t)q * ng, nx + 8 * q, nullptr, nullptr,
                              fuse ? b->counts.as<uint64_t>() : nullptr);
      hipGraph_t gr = nullptr;
      hipError_t e = hipStreamEndCapture(st, &gr);
      if (s != EH_OK) { if (gr) (void)hipGraphDestroy(gr); return s; }
      if (e != hipSuccess) { eh_set_error("graph capture: %s", hipGetErrorString(e)); return EH_ERR_HIP; }
      b->graph = gr;
      EH_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_key[0] = key[0];
      b->graph_key[1] = key[1];
    }
    EH_HIP(hipEventRecord(ctx->ev0, st));
    EH_HIP(hipGraphLaunch(b->exec, st));
    EH_HIP(hipEventRecord(ctx->ev1, st));
    EH_HIP(hipEventSynchronize(ctx->ev1));
    EH_HIP(hipEventElapsedTime(&elapsed, ctx->ev0, ctx->ev1));
  } else {
    for (int q = 0; q < 2; q++)
      if (want[q]) EH_CHECK(build_plan(b, q, key[q]));
    EH_HIP(hipEventRecord(ctx->ev0, st));
    k_zero64<<<grid_for(2ull * ng), 256, 0, st>>>(cnt, 2 * ng);
    k_zero_err<<<1, 64, 0, st>>>(nx);
    for (uint32_t g = 0; g < ng; g++) {
      for (int q = 0; q < 2; q++) {
        if (!want[q] || (fuse && q == 0)) continue;
        const eh_batch::Plan &P = b->plan[q];
        EH_CHECK(launch_parts(M, kk[q], key[q].sched, key[q].tp, P.small_rows.as<uint32_t>() + P.sr_off[g],
                              P.sr_off[g + 1] - P.sr_off[g], P.tasks_big.as<uint4>() + P.tb_off[g],
                              P.tb_off[g + 1] - P.tb_off[g], P.tasks_small.as<uint4>() + P.ts_off[g],
                              P.ts_off[g + 1] - P.ts_off[g], key[q].lds_small, key[q].lds_big,
                              b->counts.as<uint64_t>() + (uint64_t)q * ng, nx + 8 * q, nullptr, nullptr, 0, nullptr,
                              0, nullptr, nullptr, P.local_rows[0].as<uint32_t>() + P.lr_off[0][g],
                              P.lr_off[0][g + 1] - P.lr_off[0][g], P.local_rows[1].as<uint32_t>() + P.lr_off[1][g],
                              P.lr_off[1][g + 1] - P.lr_off[1][g], nullptr, nullptr,
                              fuse ? b->counts.as<uint64_t>() : nullptr));
      }
    }
    EH_HIP(hipEventRecord(ctx->ev1, st));
    EH_HIP(hipEventSynchronize(ctx->ev1));
    EH_HIP(hipEventElapsedTime(&elapsed, ctx->ev0, ctx->ev1));
  }
  uint32_t nxh[16];
  EH_HIP(hipMemcpy(nxh, nx, sizeof(nxh), hipMemcpyDeviceToHost));
  if (nxh[7] | nxh[15]) {
    eh_set_error("batch count: an LDS hash table had no free slot (device error words %#x %#x); the counts "
                 "are not valid", nxh[7], nxh[15]);
    return EH_ERR_INVALID;
  }
  std::vector<uint64_t> h(2ull * ng);
  EH_HIP(hipMemcpy(h.data(), b->counts.p, 16ull * ng, hipMemcpyDeviceToHost));
  for (uint32_t g = 0; g < ng; g++) {
    if (tri) tri[g] = h[g];
    if (k4) k4[g] = h[ng + g];
  }
  if (ms) *ms = elapsed;
  return EH_OK;
}

extern "C" eh_status eh_batch_alg_bytes(eh_batch *b, uint64_t *tri_bytes, uint64_t *k4_bytes) {
  if (!b || !b->M || !tri_bytes || !k4_bytes) { eh_set_error("eh_batch_alg_bytes: null argument"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(b->ctx->device));
  const eh_matrix *M = b->M;
  // per graph B_TC = 4 (m_g + W_g) + 8 (n_g + 1); the arena's m and W are the sums over graphs
  const uint64_t tri = 4ull * (M->card + M->wedges) + 8ull * ((uint64_t)b->nv_total + b->ng);
  uint64_t tw = 0;
  EH_CHECK(matrix_triangle_weight(b->M, b->ng, &tw));
  *tri_bytes = tri;
  *k4_bytes = tri + 4ull * tw;
  return EH_OK;
}

extern "C" void eh_batch_free(eh_batch *b) { delete b; }
