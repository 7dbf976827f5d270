// eh_mlp.hip -- the learned per-row selector.
//
// Every row that is the B side of a pair can be stored and searched in one of four arms
// (RowDesc type values), and the tuned triangle pass runs a different algorithm for each:
//   0 BS  bitset: a short source row's wedge queries probe one bit (intersection.hpp:1023-1058);
//         a task pair probes B's words with A's ids or ANDs them with A's LDS bitmap
//   1 PS  pshort: queries walk the partitions (intersection.hpp:1062-1173); pairs stream
//         B's canonical list against A's LDS structure
//   2 U   uint: queries binary-search the list; pairs stream it
//   3 UG  uint searched: queries binary-search the list; pairs search each of A's ids
//         within B's first and last id in B's list instead of streaming B (the per-pair
//         algorithm for long rows few of A's ids fall in; set_intersect_galloping's role,
//         :393-584)
// An FP32 MLP 16 -> 64 -> 64 -> 4 (ReLU) maps 16 row features to the predicted cost of each
// arm against uint (signed log2 of ns): the row's share of the change in the tuned pass's time
// (every pair A-side) and in the build time when the row's group of similar rows takes the arm,
// both measured on the engine (simdgraphprocessing_amd/selector.py, label_graph).  The LEARNED
// layout policy gives each candidate row (eh_selector_set_schedule) the arm with the smallest
// prediction among bitset, uint and uint searched: in the tuned pass a pshort B row's pairs stream
// its canonical list, as a uint row's do, and only its short-row queries differ -- too little to
// learn from timings (its labels are within the noise of uint's), and a network output that left
// its training range on graphs larger than the training graphs (RMAT-22 and up), so the pshort arm
// stays for the fixed-arm ablations and the typed builds only.  Arms change only how a count is
// computed, never which elements are common, so counts are identical under every selection.
//
// Kernels here: uses and features, the selection (MLP forward with the parameters staged in
// LDS, over the rows whose arm can matter), MLP loss + per-sample deltas, deterministic
// per-parameter gradient sums, Adam.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>

#include "eh_internal.hpp"

namespace {

constexpr int NI = EH_MLP_IN, NH = EH_MLP_HIDDEN, NO = EH_MLP_OUT;
constexpr int OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NO * NH, NPARAM = OB3 + NO;
static_assert(NPARAM == EH_MLP_PARAMS, "parameter count");

inline unsigned grid_for(uint64_t work, unsigned block = 256) {
  uint64_t g = ceil_div(work ? work : 1, block);
  return (unsigned)(g > 65536 ? 65536 : g);
}

// ---------------------------------------------------------------- uses and features

// uses[j] = rows holding j, one atomic per entry (any matrix)
__global__ void k_uses(const uint32_t *card, const uint32_t *aoff, const uint32_t *unbr, uint32_t n, uint32_t *uses) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = wid; i < n; i += nw) {
    const uint32_t c = card[i];
    const uint32_t *r = unbr + (uint64_t)aoff[i] * 4;
    for (uint32_t k = lane; k < c; k += 64) atomicAdd(&uses[r[k]], 1u);
  }
}

// uses of the lower-triangle rows of a symmetric graph: row(j) holds j's neighbours below j, so
// the rows holding j are its neighbours above j, degree(j) - card(j) of them
__global__ void k_uses_sym(const uint64_t *goff, const uint32_t *card, uint32_t n, uint32_t *uses) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    uses[i] = (uint32_t)(goff[i + 1] - goff[i]) - card[i];
}

// uses[j] under a tuned triangle schedule (split, bsh): the rows holding j whose pair (i, j) the
// pass runs A-side or edge-centric -- j at or above the split in a row that holds some B row
// there of more than bsh ids (k_row_work's test; the other rows' pairs are B-stationary and
// read canonical lists).  One wave per row; uses must start at zero.
__global__ void __launch_bounds__(256) k_uses_tuned(const uint32_t *card, const uint32_t *aoff, const uint32_t *unbr,
                                                    uint32_t n, uint32_t split, uint32_t bsh, uint32_t *uses) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = wid; i < n; i += nw) {
    const uint32_t c = card[i];
    if (c == 0) continue;
    const uint32_t *r = unbr + (uint64_t)aoff[i] * 4;
    if (r[c - 1] < split) continue;  // every pair below the split
    uint32_t bad = 0;
    for (uint32_t k = lane; k < c; k += 64) {
      const uint32_t x = r[k];
      if (x >= split && card[x] > bsh) bad = 1u;
    }
    if (__ballot(bad != 0u) == 0ull) continue;
    for (uint32_t k = lane; k < c; k += 64) {
      const uint32_t x = r[k];
      if (x >= split && card[x]) atomicAdd(&uses[x], 1u);
    }
  }
}

__device__ __forceinline__ float lg1(float x) { return log2f(1.0f + x); }

// the statistics behind a row's features
struct RowStats {
  uint32_t c, lo, hi, mid, parts, low, quarter, maxgap;  // low: ids below n / 64
};

// 16 features of row i, all in [0, 1] and free of the graph's size where they can be (ids and
// spans relative to n, gaps relative to the span, counts relative to the row), so a network
// trained on small graphs reads a large graph's rows on the same scale.  The absolute sizes
// saturate where the training graphs (RMAT 18-21) end -- a length of FEAT_CARD_CAP, FEAT_USES_CAP
// uses -- so a larger graph's heaviest rows read as the heaviest rows the network has seen
// instead of outside its range.  In a batch arena, ids, i and n are those of the row's own
// graph (ids less the graph's first vertex), so an arena row reads as the same row of its graph
// alone; feature 15 tells the two regimes apart (the arena's rows are read by the batch passes'
// edge-centric wedge queries, a single graph's by its tuned pass):
//   0 log length            1 log span / n         2 density (ids per id of span)
//   3 16-bit partitions per id   4 log uses        5 log (uses x length)
//   6 log bitset words per id    7 share of ids below n / 64
//   8 log span per id       9 largest gap / span   10 row id / n
//   11 first id / n         12 last id / n         13 middle id / n
//   14 share of ids in the first quarter of the span   15 1 (one graph) or 0 (batch arena)
constexpr float FEAT_CARD_CAP = 1023.0f, FEAT_USES_CAP = 16383.0f;
__device__ __forceinline__ void write_features(const RowStats &r, uint32_t u_raw, uint32_t i, uint32_t n, float flag,
                                               float *out) {
  float f[NI];
  for (int q = 0; q < NI; q++) f[q] = 0.0f;
  const float u = fminf((float)u_raw, FEAT_USES_CAP), fn = (float)n;
  if (r.c) {
    const float span = (float)(r.hi - r.lo), c = (float)r.c;
    const float words = (float)((r.hi >> 6) - (r.lo >> 6) + 1u);
    f[0] = lg1(fminf(c, FEAT_CARD_CAP)) / 16.0f;
    f[1] = lg1(1048576.0f * span / fn) / 20.0f;
    f[2] = c / (span + 1.0f);
    f[3] = (float)r.parts / c;
    f[6] = lg1(words / c) / 16.0f;
    f[7] = (float)r.low / c;
    f[8] = lg1(span / c) / 32.0f;
    f[9] = span > 0.0f ? (float)r.maxgap / span : 0.0f;
    f[11] = (float)r.lo / fn;
    f[12] = (float)r.hi / fn;
    f[13] = (float)r.mid / fn;
    f[14] = (float)r.quarter / c;
  }
  f[4] = lg1(u) / 24.0f;
  f[5] = lg1(u * fminf((float)r.c, FEAT_CARD_CAP)) / 40.0f;
  f[10] = (float)i / fn;
  f[15] = flag;
  float4 *o = reinterpret_cast<float4 *>(out);
#pragma unroll
  for (int q = 0; q < NI / 4; q++) o[q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
}

// the same 16 features into an LDS queue entry (word stores: entries are an odd stride apart)
__device__ __forceinline__ void write_features_lds(const RowStats &r, uint32_t u, uint32_t i, uint32_t n, float flag,
                                                   float *out) {
  float f[NI];
  write_features(r, u, i, n, flag, f);
#pragma unroll
  for (int q = 0; q < NI; q++) out[q] = f[q];
}

__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr uint32_t FEAT_LANE_MAX = 64;  // rows up to this long are read by one lane

// the statistics of a row of at most FEAT_LANE_MAX ids, read by one lane; ids are taken less
// `base` (the first vertex of the row's graph in a batch arena, else 0).  The row is read as
// 16-byte vectors (rows start on 16 bytes and are padded to four ids), 32 ids at a time with
// every load of a round issued before any id is used
__device__ __forceinline__ RowStats lane_stats(const uint32_t *r, uint32_t c, uint32_t lowcut, uint32_t base) {
  RowStats st = {c, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
  if (!c) return st;
  st.hi = r[c - 1] - base;
  st.mid = r[c / 2] - base;
  st.lo = r[0] - base;
  st.parts = 1;
  const uint32_t qcut = st.lo + (st.hi - st.lo) / 4u;
  uint32_t prev = st.lo;
  for (uint32_t k0 = 0; k0 < c; k0 += 32u) {
    uint4 v[8];
#pragma unroll
    for (uint32_t q = 0; q < 8u; q++)
      v[q] = (k0 + 4u * q < c) ? reinterpret_cast<const uint4 *>(r + k0)[q] : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (uint32_t k = 0; k < 32u; k++) {
      if (k0 + k < c) {
        const uint4 w = v[k >> 2];
        const uint32_t x = ((k & 3u) == 0u ? w.x : (k & 3u) == 1u ? w.y : (k & 3u) == 2u ? w.z : w.w) - base;
        if (k0 + k) {
          st.parts += (x >> 16) != (prev >> 16);
          st.maxgap = max(st.maxgap, x - prev);
        }
        st.low += x < lowcut;
        st.quarter += x <= qcut;
        prev = x;
      }
    }
  }
  return st;
}

// the statistics of a longer row, read by the whole wave (every lane returns them)
__device__ __forceinline__ RowStats wave_stats(const uint32_t *r, uint32_t c, uint32_t lane, uint32_t lowcut,
                                               uint32_t base) {
  RowStats st = {c, r[0] - base, r[c - 1] - base, r[c / 2] - base, 0u, 0u, 0u, 0u};
  const uint32_t qcut = st.lo + (st.hi - st.lo) / 4u;
  uint32_t parts = 0, low = 0, quart = 0, gap = 0;
  for (uint32_t k = lane; k < c; k += 64) {
    const uint32_t x = r[k] - base;
    if (k) {
      const uint32_t pv = r[k - 1] - base;
      parts += (x >> 16) != (pv >> 16);
      gap = max(gap, x - pv);
    }
    low += x < lowcut;
    quart += x <= qcut;
  }
  for (int o = 32; o > 0; o >>= 1) {
    parts += __shfl_xor(parts, o, 64);
    low += __shfl_xor(low, o, 64);
    quart += __shfl_xor(quart, o, 64);
    gap = max(gap, (uint32_t)__shfl_xor(gap, o, 64));
  }
  st.parts = 1u + parts;
  st.low = low;
  st.quarter = quart;
  st.maxgap = gap;
  return st;
}

// spans[g] = {first vertex, end} of graph g of a batch arena (graph_of: the graph of every row,
// ascending, every graph's rows contiguous)
__global__ void k_graph_spans(const uint32_t *graph_of, uint32_t n, uint2 *spans) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t g = graph_of[i];
    if (i == 0 || graph_of[i - 1] != g) spans[g].x = (uint32_t)i;
    if (i + 1 == n || graph_of[i + 1] != g) spans[g].y = (uint32_t)i + 1u;
  }
}

// the graph a row's ids and index are taken in: its own graph in a batch arena (graph_of set),
// else the whole matrix
struct RowGraph {
  uint32_t base, n;
};
__device__ __forceinline__ RowGraph row_graph(const uint32_t *graph_of, const uint2 *spans, uint64_t i, uint32_t n) {
  if (!graph_of) return RowGraph{0u, n};
  const uint2 sp = spans[graph_of[i]];
  return RowGraph{sp.x, sp.y - sp.x};
}

// Features of every row, at its own index.  A wave takes 64 consecutive rows: each lane reads
// a row of at most FEAT_LANE_MAX ids itself, then the wave reads the longer ones together.
__global__ void __launch_bounds__(256) k_features(const uint32_t *card, const uint32_t *aoff, const uint32_t *unbr,
                                                  const uint32_t *uses, uint32_t n, const uint32_t *graph_of,
                                                  const uint2 *spans, float *F) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwv = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const float flag = graph_of ? 0.0f : 1.0f;
  for (uint64_t base = wid * 64; base < n; base += nwv * 64) {
    const uint64_t i = base + lane;
    const bool valid = i < n;
    const uint32_t c = valid ? card[i] : 0u;
    const uint32_t u = valid ? uses[i] : 0u;
    const RowGraph rg = valid ? row_graph(graph_of, spans, i, n) : RowGraph{0u, n};
    if (valid && c <= FEAT_LANE_MAX)
      write_features(lane_stats(unbr + (uint64_t)aoff[i] * 4, c, rg.n >> 6, rg.base), u, (uint32_t)i - rg.base, rg.n,
                     flag, F + i * NI);
    uint64_t lm = __ballot(valid && c > FEAT_LANE_MAX);
    while (lm) {
      const uint32_t q = (uint32_t)__builtin_ctzll(lm);
      lm &= lm - 1;
      const uint32_t ii = (uint32_t)base + q;
      const uint32_t cc = __builtin_amdgcn_readlane(c, q), uu = __builtin_amdgcn_readlane(u, q);
      const uint32_t gb = __builtin_amdgcn_readlane(rg.base, q), gn = __builtin_amdgcn_readlane(rg.n, q);
      const RowStats st = wave_stats(unbr + (uint64_t)aoff[ii] * 4, cc, lane, gn >> 6, gb);
      if (lane == 0) write_features(st, uu, ii - gb, gn, flag, F + (uint64_t)ii * NI);
    }
  }
}

// ---------------------------------------------------------------- MLP forward and selection

__device__ __forceinline__ float relu(float v) { return v > 0.0f ? v : 0.0f; }

// y = MLP(x) with the parameters W (LDS: every lane of a wave reads the same word)
__device__ __forceinline__ void mlp_eval(const float *W, const float *__restrict__ X, float (&y)[NO]) {
  float x[NI], h1[NH];
  const float4 *X4 = reinterpret_cast<const float4 *>(X);
#pragma unroll
  for (int q = 0; q < NI / 4; q++) {
    const float4 v = X4[q];
    x[4 * q] = v.x;
    x[4 * q + 1] = v.y;
    x[4 * q + 2] = v.z;
    x[4 * q + 3] = v.w;
  }
#pragma unroll
  for (int o = 0; o < NH; o++) {
    float a = W[OB1 + o];
#pragma unroll
    for (int i = 0; i < NI; i++) a += W[OW1 + o * NI + i] * x[i];
    h1[o] = relu(a);
  }
#pragma unroll
  for (int k = 0; k < NO; k++) y[k] = W[OB3 + k];
  for (int o = 0; o < NH; o++) {
    float a = W[OB2 + o];
#pragma unroll
    for (int i = 0; i < NH; i++) a += W[OW2 + o * NH + i] * h1[i];
    const float h2 = relu(a);
#pragma unroll
    for (int k = 0; k < NO; k++) y[k] += W[OW3 + k * NH + o] * h2;
  }
}

// the block stages the parameters in LDS once (22 KB), then one thread per sample
__global__ void __launch_bounds__(256) k_mlp_forward(const float *__restrict__ P, const float *__restrict__ X,
                                                     int batch, float *Y) {
  __shared__ float W[NPARAM];
  for (int q = threadIdx.x; q < NPARAM; q += blockDim.x) W[q] = P[q];
  __syncthreads();
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < batch; b += gridDim.x * blockDim.x) {
    float y[NO];
    mlp_eval(W, X + (size_t)b * NI, y);
    for (int k = 0; k < NO; k++) Y[(size_t)b * NO + k] = y[k];
  }
}

// The layout of every row under the LEARNED policy, in one pass.  Rows of fewer than two ids,
// or held by no row, are uint (their arm cannot matter).  The others ("candidates") take the
// arm with the smallest predicted cost (fixed < 0) or arm `fixed` (the ablations); a bitset
// only where its words stay within `guard` per id.
// Each wave walks 64 consecutive rows at a time and appends its candidates' features to its own
// LDS queue (short rows read by one lane, long rows by the wave, as in k_features); whenever 64
// are queued, the wave evaluates the network for all 64 at once on the matrix cores
// (mlp_mfma64) with the parameters the block staged in LDS, then the queue shifts down.  No
// features leave the CU and no candidate list is built in HBM.
constexpr uint32_t SEL_Q = 128;        // queue entries per wave (below 64 + one step's 64)
constexpr uint32_t SEL_STRIDE = NI + 1;  // LDS words per entry (odd: lanes read distinct banks)
static_assert(NI == 16 && NH == 64 && NO == 4, "the matrix-core evaluation's tiling");

// The network over a wave's 64 queued rows on the matrix cores (v_mfma_f32_16x16x4_f32, FP32
// in and out), in transposed form so that no activation crosses lanes: H1^T = W1 X^T, H2^T =
// W2 relu(H1^T), Y^T = W3 relu(H2^T), each a sum of 16x16x4 products.  In a product's result
// lane l (q = l / 16, c = l % 16) holds rows 4q .. 4q + 3 of column c; the next layer takes
// those four rows as the k-slot q of four products, so its B operand is the previous result's
// register as it stands (k-slot q of the product for register r of hidden tile hb is hidden
// unit 16 hb + 4q + r), and only the weights are read in a matching order.  The block stages the
// weights in that order (mlp_mfma_order: 96 A operands x 64 lanes, then b1, b2, b3), so every
// operand is one conflict-free LDS word per lane.  Rows of X^T's columns are the queued rows
// (x = 16 xb + c), so a queue slot past the rows being drained only affects its own column.
// y[xb] of lanes 0..15 returns the 4 outputs of row 16 xb + lane.
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr uint32_t MM_A1 = 0, MM_A2 = 16, MM_A3 = 80, MM_OPS = 96, MM_B1 = MM_OPS * 64, MM_B2 = MM_B1 + NH,
                   MM_B3 = MM_B2 + NH, MM_WORDS = MM_B3 + NO;

__device__ __forceinline__ void mlp_mfma64(const float *Wm, const float *QF, uint32_t lane, v4f (&y)[4]) {
  const uint32_t q = lane >> 4, c = lane & 15u;
  v4f h1[4][4];  // [hidden tile][row tile]
#pragma unroll
  for (int hb = 0; hb < 4; hb++)
#pragma unroll
    for (int xb = 0; xb < 4; xb++)
#pragma unroll
      for (int r = 0; r < 4; r++) h1[hb][xb][r] = Wm[MM_B1 + hb * 16 + 4 * q + r];
#pragma unroll
  for (int kc = 0; kc < NI / 4; kc++) {
    float b[4];
#pragma unroll
    for (int xb = 0; xb < 4; xb++) b[xb] = QF[(xb * 16 + c) * SEL_STRIDE + 4 * kc + q];
#pragma unroll
    for (int hb = 0; hb < 4; hb++) {
      const float a = Wm[(MM_A1 + hb * 4 + kc) * 64 + lane];
#pragma unroll
      for (int xb = 0; xb < 4; xb++) h1[hb][xb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[xb], h1[hb][xb], 0, 0, 0);
    }
  }
#pragma unroll
  for (int hb = 0; hb < 4; hb++)
#pragma unroll
    for (int xb = 0; xb < 4; xb++)
#pragma unroll
      for (int r = 0; r < 4; r++) h1[hb][xb][r] = relu(h1[hb][xb][r]);
#pragma unroll
  for (int xb = 0; xb < 4; xb++)
#pragma unroll
    for (int r = 0; r < 4; r++) y[xb][r] = q == 0 ? Wm[MM_B3 + r] : 0.0f;
#pragma unroll
  for (int ob = 0; ob < 4; ob++) {
    v4f h2[4];
#pragma unroll
    for (int xb = 0; xb < 4; xb++)
#pragma unroll
      for (int r = 0; r < 4; r++) h2[xb][r] = Wm[MM_B2 + ob * 16 + 4 * q + r];
#pragma unroll
    for (int hb = 0; hb < 4; hb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const float a = Wm[(MM_A2 + (ob * 4 + hb) * 4 + r) * 64 + lane];
#pragma unroll
        for (int xb = 0; xb < 4; xb++) h2[xb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, h1[hb][xb][r], h2[xb], 0, 0, 0);
      }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const float a = Wm[(MM_A3 + ob * 4 + r) * 64 + lane];
#pragma unroll
      for (int xb = 0; xb < 4; xb++) y[xb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, relu(h2[xb][r]), y[xb], 0, 0, 0);
    }
  }
}

// the arm with the smallest prediction among bitset (within the guard), uint and uint searched
__device__ __forceinline__ uint8_t arm_of(const v4f &y, uint32_t c, uint32_t nw, uint32_t guard) {
  const bool bs_ok = (uint64_t)nw <= (uint64_t)guard * c;
  uint8_t pick = EH_U;
  float best = 3.0e38f;
#pragma unroll
  for (uint32_t k = 0; k < NO; k++) {
    if (k == EH_PS || (k == EH_BS && !bs_ok)) continue;
    if (y[k] < best) { best = y[k]; pick = (uint8_t)k; }
  }
  return pick;
}

__global__ void __launch_bounds__(256) k_select_learned(const uint32_t *card, const uint32_t *aoff,
                                                        const uint32_t *unbr, const uint32_t *uses, uint32_t n,
                                                        const uint32_t *graph_of, const uint2 *spans,
                                                        const float *__restrict__ P, uint32_t guard, int fixed,
                                                        uint8_t *types) {
  __shared__ float Wm[MM_WORDS];
  __shared__ float Qf[4][SEL_Q * SEL_STRIDE];
  __shared__ uint32_t Qr[4][SEL_Q], Qc[4][SEL_Q], Qw[4][SEL_Q];
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  if (fixed < 0) {
    for (uint32_t q = threadIdx.x; q < MM_WORDS; q += blockDim.x) Wm[q] = P[q];
  }
  __syncthreads();
  float *QF = Qf[wv];
  uint32_t *QR = Qr[wv], *QC = Qc[wv], *QW = Qw[wv];
  uint32_t qn = 0;
  const float flag = graph_of ? 0.0f : 1.0f;
  auto drain = [&](uint32_t cnt) {
    // the first cnt queued rows (row 16 xb + lane's outputs in lane `lane` < 16)
    v4f y[4];
    mlp_mfma64(Wm, QF, lane, y);
#pragma unroll
    for (uint32_t xb = 0; xb < 4; xb++) {
      const uint32_t e = xb * 16u + lane;
      if (lane < 16u && e < cnt) types[QR[e]] = arm_of(y[xb], QC[e], QW[e], guard);
    }
    wave_sync();
    const uint32_t rest = qn - cnt;
    for (uint32_t e = lane; e < rest; e += 64) {
      for (uint32_t q = 0; q < NI; q++) QF[e * SEL_STRIDE + q] = QF[(e + cnt) * SEL_STRIDE + q];
      QR[e] = QR[e + cnt];
      QC[e] = QC[e + cnt];
      QW[e] = QW[e + cnt];
    }
    wave_sync();
    qn = rest;
  };
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwv = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t base = wid * 64; base < n; base += nwv * 64) {
    const uint64_t i = base + lane;
    const bool valid = i < n;
    const uint32_t c = valid ? card[i] : 0u;
    const uint32_t u = valid ? uses[i] : 0u;
    const bool want = valid && c >= 2u && u >= 1u;
    const uint32_t *r = unbr + (uint64_t)(valid ? aoff[i] : 0u) * 4;
    if (valid && !want) types[i] = EH_U;
    if (fixed >= 0) {
      if (want) {
        const uint32_t nw = (r[c - 1] >> 6) - (r[0] >> 6) + 1u;
        types[i] = (fixed == EH_BS && (uint64_t)nw > (uint64_t)guard * c) ? (uint8_t)EH_U : (uint8_t)fixed;
      }
      continue;
    }
    const uint64_t m = __ballot(want);
    const uint32_t slot = qn + mbcnt64(m);
    const RowGraph rg = want ? row_graph(graph_of, spans, i, n) : RowGraph{0u, n};
    if (want) {
      QR[slot] = (uint32_t)i;
      QC[slot] = c;
      QW[slot] = (r[c - 1] >> 6) - (r[0] >> 6) + 1u;
      if (c <= FEAT_LANE_MAX)
        write_features_lds(lane_stats(r, c, rg.n >> 6, rg.base), u, (uint32_t)i - rg.base, rg.n, flag,
                           QF + slot * SEL_STRIDE);
    }
    uint64_t lm = __ballot(want && c > FEAT_LANE_MAX);
    while (lm) {
      const uint32_t q = (uint32_t)__builtin_ctzll(lm);
      lm &= lm - 1;
      const uint32_t cc = __builtin_amdgcn_readlane(c, q), uu = __builtin_amdgcn_readlane(u, q);
      const uint32_t sq = __builtin_amdgcn_readlane(slot, q);
      const uint32_t gb = __builtin_amdgcn_readlane(rg.base, q), gn = __builtin_amdgcn_readlane(rg.n, q);
      const uint32_t ii = (uint32_t)base + q;
      const RowStats st = wave_stats(unbr + (uint64_t)aoff[ii] * 4, cc, lane, gn >> 6, gb);
      if (lane == 0) write_features_lds(st, uu, ii - gb, gn, flag, QF + sq * SEL_STRIDE);
    }
    qn += (uint32_t)__popcll(m);
    wave_sync();
    if (qn >= 64u) drain(64u);
  }
  if (fixed < 0 && qn) drain(qn);
}

// per-sample activations and deltas of the masked mean squared error
// (loss = sum(mask * (y - t)^2) / sum(mask)); writes for sample b:
//   H1[b][64], H2[b][64], D1[b][64] (dL/dz1), D2[b][64] (dL/dz2), DY[b][4], L[b]
__global__ void k_mlp_deltas(const float *__restrict__ P, const float *__restrict__ X, const float *__restrict__ T,
                             const float *__restrict__ Mk, const int *__restrict__ idx, int batch, float inv_lab,
                             float *H1, float *H2, float *D1, float *D2, float *DY, float *L) {
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < batch; b += gridDim.x * blockDim.x) {
    const size_t sIdx = idx ? (size_t)idx[b] : (size_t)b;
    float x[NI], z1[NH], z2[NH], dy[NO];
    for (int i = 0; i < NI; i++) x[i] = X[sIdx * NI + i];
    for (int o = 0; o < NH; o++) {
      float a = P[OB1 + o];
      for (int i = 0; i < NI; i++) a += P[OW1 + o * NI + i] * x[i];
      z1[o] = a;
      H1[(size_t)b * NH + o] = relu(a);
    }
    float y[NO];
    for (int k = 0; k < NO; k++) y[k] = P[OB3 + k];
    for (int o = 0; o < NH; o++) {
      float a = P[OB2 + o];
      for (int i = 0; i < NH; i++) a += P[OW2 + o * NH + i] * relu(z1[i]);
      z2[o] = a;
      const float h2 = relu(a);
      H2[(size_t)b * NH + o] = h2;
      for (int k = 0; k < NO; k++) y[k] += P[OW3 + k * NH + o] * h2;
    }
    float loss = 0.0f;
    for (int k = 0; k < NO; k++) {
      const float m = Mk[sIdx * NO + k];
      const float d = y[k] - T[sIdx * NO + k];
      loss += m * d * d;
      dy[k] = m * 2.0f * d * inv_lab;
      DY[(size_t)b * NO + k] = dy[k];
    }
    L[b] = loss;
    float d2[NH];
    for (int o = 0; o < NH; o++) {
      float g = 0.0f;
      for (int k = 0; k < NO; k++) g += P[OW3 + k * NH + o] * dy[k];
      d2[o] = z2[o] > 0.0f ? g : 0.0f;
      D2[(size_t)b * NH + o] = d2[o];
    }
    for (int i = 0; i < NH; i++) {
      float g = 0.0f;
      for (int o = 0; o < NH; o++) g += P[OW2 + o * NH + i] * d2[o];
      D1[(size_t)b * NH + i] = z1[i] > 0.0f ? g : 0.0f;
    }
  }
}

// one thread per parameter: gradient = sum over the batch in sample order
__global__ void k_mlp_grad(const float *__restrict__ X, const int *__restrict__ idx, int batch, const float *H1,
                           const float *H2, const float *D1, const float *D2, const float *DY, const float *L,
                           float inv_lab, float *G, float *loss_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float g = 0.0f;
  if (p < OB1) {
    const int o = (p - OW1) / NI, i = (p - OW1) % NI;
    for (int b = 0; b < batch; b++) g += D1[(size_t)b * NH + o] * X[(idx ? (size_t)idx[b] : (size_t)b) * NI + i];
  } else if (p < OW2) {
    const int o = p - OB1;
    for (int b = 0; b < batch; b++) g += D1[(size_t)b * NH + o];
  } else if (p < OB2) {
    const int o = (p - OW2) / NH, i = (p - OW2) % NH;
    for (int b = 0; b < batch; b++) g += D2[(size_t)b * NH + o] * H1[(size_t)b * NH + i];
  } else if (p < OW3) {
    const int o = p - OB2;
    for (int b = 0; b < batch; b++) g += D2[(size_t)b * NH + o];
  } else if (p < OB3) {
    const int k = (p - OW3) / NH, o = (p - OW3) % NH;
    for (int b = 0; b < batch; b++) g += DY[(size_t)b * NO + k] * H2[(size_t)b * NH + o];
  } else {
    const int k = p - OB3;
    for (int b = 0; b < batch; b++) g += DY[(size_t)b * NO + k];
  }
  G[p] = g;
  if (p == 0 && loss_out) {
    float s = 0.0f;
    for (int b = 0; b < batch; b++) s += L[b];
    *loss_out = s * inv_lab;
  }
}

__global__ void k_adam(float *P, const float *G, float *M1, float *M2, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  const float g = G[p];
  const float m = b1 * M1[p] + (1.0f - b1) * g;
  const float v = b2 * M2[p] + (1.0f - b2) * g * g;
  M1[p] = m;
  M2[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace

// ---------------------------------------------------------------- host helpers

// uses of every row: from the source graph's degrees when the matrix holds the lower
// triangle of a symmetric graph (sym_off: its row offsets), else one atomic per entry
// uses of every row: under the context's selector schedule for an oriented single-graph
// matrix with the split on (k_uses_tuned), else every row that holds it.  *none: set when no
// row has a use (no row at or above the split is longer than the k_bshl bound, so every pair
// is B-stationary: then the rows are not scanned)
static eh_status uses_dev(eh_matrix *m, const uint64_t *sym_off, DevBuf &uses, bool *none = nullptr) {
  hipStream_t st = m->ctx->stream;
  const uint32_t n = m->n;
  EH_CHECK(uses.alloc(4ull * n + 4));
  const eh_ctx *ctx = m->ctx;
  if (none) *none = false;
  if (ctx->sel_split && m->filter == EH_FILTER_LOWER && !m->dgraph_of.p) {
    EH_HIP(hipMemsetAsync(uses.p, 0, 4ull * n + 4, st));
    uint32_t longest = 0;
    if (ctx->sel_split < n)
      EH_CHECK(prim_reduce_max_u32(m->ctx, m->dcard.as<uint32_t>() + ctx->sel_split, n - ctx->sel_split, &longest));
    if (longest <= ctx->sel_bsh) {
      if (none) *none = true;
      return EH_OK;
    }
    k_uses_tuned<<<grid_for((uint64_t)n * 64), 256, 0, st>>>(m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(),
                                                             m->dunbr.as<uint32_t>(), n, ctx->sel_split,
                                                             ctx->sel_bsh, uses.as<uint32_t>());
  } else if (sym_off && m->filter == EH_FILTER_LOWER) {
    k_uses_sym<<<grid_for(n), 256, 0, st>>>(sym_off, m->dcard.as<uint32_t>(), n, uses.as<uint32_t>());
  } else {
    EH_HIP(hipMemsetAsync(uses.p, 0, 4ull * n + 4, st));
    k_uses<<<grid_for((uint64_t)n * 64), 256, 0, st>>>(m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(),
                                                       m->dunbr.as<uint32_t>(), n, uses.as<uint32_t>());
  }
  EH_HIP(hipGetLastError());
  return EH_OK;
}

// {first vertex, end} of every graph of a batch arena; spans stays empty for a single graph
static eh_status graph_spans_dev(eh_matrix *m, DevBuf &spans) {
  if (!m->dgraph_of.p || m->n == 0) return EH_OK;
  hipStream_t st = m->ctx->stream;
  uint32_t last = 0;
  EH_HIP(hipMemcpyAsync(&last, m->dgraph_of.as<uint32_t>() + (m->n - 1), 4, hipMemcpyDeviceToHost, st));
  EH_HIP(hipStreamSynchronize(st));
  EH_CHECK(spans.alloc(8ull * ((uint64_t)last + 1) + 8));
  k_graph_spans<<<grid_for(m->n), 256, 0, st>>>(m->dgraph_of.as<uint32_t>(), m->n, spans.as<uint2>());
  EH_HIP(hipGetLastError());
  return EH_OK;
}

// features of every row, at its own index
static eh_status features_dev(eh_matrix *m, DevBuf &F, DevBuf &uses, DevBuf &spans) {
  hipStream_t st = m->ctx->stream;
  const uint32_t n = m->n;
  EH_CHECK(uses_dev(m, nullptr, uses));
  EH_CHECK(graph_spans_dev(m, spans));
  EH_CHECK(F.alloc(4ull * NI * n + 64));
  k_features<<<grid_for((uint64_t)n, 256), 256, 0, st>>>(m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(),
                                                         m->dunbr.as<uint32_t>(), uses.as<uint32_t>(), n,
                                                         spans.p ? m->dgraph_of.as<uint32_t>() : nullptr,
                                                         spans.as<uint2>(), F.as<float>());
  EH_HIP(hipGetLastError());
  return EH_OK;
}

static eh_status mlp_forward_dev(eh_ctx *ctx, const float *d_params, const float *d_x, int batch, float *d_y) {
  k_mlp_forward<<<grid_for((uint64_t)batch, 256), 256, 0, ctx->stream>>>(d_params, d_x, batch, d_y);
  EH_HIP(hipGetLastError());
  return EH_OK;
}

// the parameters in the order mlp_mfma64 reads them: for each of its 96 A operands the 64
// lanes' values (lane l: q = l / 16, c = l % 16) -- W1 per (hidden tile hb, k chunk kc):
// W1[16 hb + c][4 kc + q]; W2 per (out tile ob, hidden tile hb, register r): W2[16 ob + c][16 hb +
// 4 q + r]; W3 per (ob, r): W3[c][16 ob + 4 q + r] for c < 4, else 0 -- then b1, b2, b3
static void mlp_mfma_order(const float *P, std::vector<float> &wm) {
  wm.assign(MM_WORDS, 0.0f);
  for (uint32_t lane = 0; lane < 64; lane++) {
    const uint32_t q = lane >> 4, c = lane & 15u;
    for (uint32_t hb = 0; hb < 4; hb++)
      for (uint32_t kc = 0; kc < NI / 4; kc++)
        wm[(MM_A1 + hb * 4 + kc) * 64 + lane] = P[OW1 + (hb * 16 + c) * NI + 4 * kc + q];
    for (uint32_t ob = 0; ob < 4; ob++)
      for (uint32_t hb = 0; hb < 4; hb++)
        for (uint32_t r = 0; r < 4; r++)
          wm[(MM_A2 + (ob * 4 + hb) * 4 + r) * 64 + lane] = P[OW2 + (ob * 16 + c) * NH + hb * 16 + 4 * q + r];
    for (uint32_t ob = 0; ob < 4; ob++)
      for (uint32_t r = 0; r < 4; r++)
        wm[(MM_A3 + ob * 4 + r) * 64 + lane] = c < (uint32_t)NO ? P[OW3 + c * NH + ob * 16 + 4 * q + r] : 0.0f;
  }
  for (int o = 0; o < NH; o++) wm[MM_B1 + o] = P[OB1 + o];
  for (int o = 0; o < NH; o++) wm[MM_B2 + o] = P[OB2 + o];
  for (int k = 0; k < NO; k++) wm[MM_B3 + k] = P[OB3 + k];
}

eh_status matrix_select_learned(eh_matrix *m, uint8_t *d_types, const uint64_t *sym_off) {
  eh_ctx *ctx = m->ctx;
  const bool fixed = ctx->sel_fixed >= 0;
  if (!fixed && ctx->mlp.size() != (size_t)NPARAM) {
    eh_set_error("EH_LAYOUT_LEARNED needs trained selector parameters (eh_selector_set_params / eh_selector_train)");
    return EH_ERR_INVALID;
  }
  hipStream_t st = ctx->stream;
  const uint32_t n = m->n;
  DevBuf uses;
  bool none = false;
  EH_CHECK(uses_dev(m, sym_off, uses, &none));
  if (none) {  // no candidate: every row uint
    EH_HIP(hipMemsetAsync(d_types, EH_U, n, st));
    EH_HIP(hipStreamSynchronize(st));
    return EH_OK;
  }
  std::vector<float> wm;
  if (!fixed) {
    mlp_mfma_order(ctx->mlp.data(), wm);
    if (ctx->mlp_dev.bytes < 4ull * MM_WORDS) EH_CHECK(ctx->mlp_dev.alloc(4ull * MM_WORDS));
    EH_HIP(hipMemcpyAsync(ctx->mlp_dev.p, wm.data(), 4ull * MM_WORDS, hipMemcpyHostToDevice, st));
  }
  // every block stages the parameters (22 KB, from L2) once: a few thousand blocks at most
  const unsigned blocks = std::min(grid_for(n), 8192u);
  DevBuf spans;
  EH_CHECK(graph_spans_dev(m, spans));
  k_select_learned<<<blocks, 256, 0, st>>>(m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(), m->dunbr.as<uint32_t>(),
                                           uses.as<uint32_t>(), n, spans.p ? m->dgraph_of.as<uint32_t>() : nullptr,
                                           spans.as<uint2>(), fixed ? nullptr : ctx->mlp_dev.as<float>(),
                                           ctx->sel_guard, ctx->sel_fixed, d_types);
  EH_HIP(hipGetLastError());
  EH_HIP(hipStreamSynchronize(st));
  return EH_OK;
}

// ---------------------------------------------------------------- C ABI

extern "C" eh_status eh_selector_set_params(eh_ctx *ctx, const float *params) {
  if (!ctx || !params) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  ctx->mlp.assign(params, params + NPARAM);
  return EH_OK;
}

extern "C" eh_status eh_selector_get_params(eh_ctx *ctx, float *params) {
  if (!ctx || !params) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  if (ctx->mlp.size() != (size_t)NPARAM) { eh_set_error("no selector parameters set"); return EH_ERR_INVALID; }
  std::copy(ctx->mlp.begin(), ctx->mlp.end(), params);
  return EH_OK;
}

extern "C" eh_status eh_selector_set_rule(eh_ctx *ctx, uint32_t guard_words, int32_t fixed_arm) {
  if (!ctx) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  if (guard_words == 0 || guard_words > (1u << 20)) { eh_set_error("bitset guard must be 1 .. 2^20 words per id"); return EH_ERR_INVALID; }
  if (fixed_arm < -1 || fixed_arm > 3) { eh_set_error("fixed arm must be -1 (learned) or 0 .. 3"); return EH_ERR_INVALID; }
  ctx->sel_guard = guard_words;
  ctx->sel_fixed = fixed_arm;
  return EH_OK;
}

extern "C" eh_status eh_selector_set_schedule(eh_ctx *ctx, int32_t bstat_split, int32_t bsh_max) {
  if (!ctx) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  if (bstat_split > 65536 || (bstat_split > 0 && bstat_split < 64)) {
    eh_set_error("bstat_split must be 0 (default), < 0 (off) or 64 .. 65536");
    return EH_ERR_INVALID;
  }
  if (bsh_max > 512) { eh_set_error("bsh_max must be at most 512"); return EH_ERR_INVALID; }
  ctx->sel_split = bstat_split < 0 ? 0u : bstat_split == 0 ? 65536u : (uint32_t)bstat_split;
  ctx->sel_bsh = bsh_max < 0 ? 0u : bsh_max == 0 ? 512u : (uint32_t)bsh_max;
  return EH_OK;
}

extern "C" eh_status eh_matrix_features(eh_matrix *m, float *features) {
  if (!m || !features) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(m->ctx->device));
  DevBuf F, uses, spans;
  EH_CHECK(features_dev(m, F, uses, spans));
  EH_HIP(hipStreamSynchronize(m->ctx->stream));
  EH_HIP(hipMemcpy(features, F.p, 4ull * NI * m->n, hipMemcpyDeviceToHost));
  return EH_OK;
}

extern "C" eh_status eh_mlp_forward(eh_ctx *ctx, const float *params, const float *x, int batch, float *y) {
  if (!ctx || !params || !x || !y || batch <= 0) { eh_set_error("bad argument"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(ctx->device));
  DevBuf P, X, Y;
  EH_CHECK(P.alloc(4ull * NPARAM));
  EH_CHECK(X.alloc(4ull * NI * batch));
  EH_CHECK(Y.alloc(4ull * NO * batch));
  EH_HIP(hipMemcpy(P.p, params, 4ull * NPARAM, hipMemcpyHostToDevice));
  EH_HIP(hipMemcpy(X.p, x, 4ull * NI * batch, hipMemcpyHostToDevice));
  EH_CHECK(mlp_forward_dev(ctx, P.as<float>(), X.as<float>(), batch, Y.as<float>()));
  EH_HIP(hipStreamSynchronize(ctx->stream));
  EH_HIP(hipMemcpy(y, Y.p, 4ull * NO * batch, hipMemcpyDeviceToHost));
  return EH_OK;
}

namespace {

struct TrainBufs {
  DevBuf H1, H2, D1, D2, DY, L, G, loss;
  eh_status alloc(int batch) {
    EH_CHECK(H1.alloc(4ull * NH * batch));
    EH_CHECK(H2.alloc(4ull * NH * batch));
    EH_CHECK(D1.alloc(4ull * NH * batch));
    EH_CHECK(D2.alloc(4ull * NH * batch));
    EH_CHECK(DY.alloc(4ull * NO * batch));
    EH_CHECK(L.alloc(4ull * batch));
    EH_CHECK(G.alloc(4ull * NPARAM));
    EH_CHECK(loss.alloc(64));
    return EH_OK;
  }
};

eh_status grad_dev(eh_ctx *ctx, const float *P, const float *X, const float *T, const float *Mk, const int *idx,
                   int batch, float inv_lab, TrainBufs &B) {
  hipStream_t st = ctx->stream;
  k_mlp_deltas<<<grid_for((uint64_t)batch, 64), 64, 0, st>>>(P, X, T, Mk, idx, batch, inv_lab, B.H1.as<float>(),
                                                             B.H2.as<float>(), B.D1.as<float>(), B.D2.as<float>(),
                                                             B.DY.as<float>(), B.L.as<float>());
  k_mlp_grad<<<grid_for(NPARAM, 256), 256, 0, st>>>(X, idx, batch, B.H1.as<float>(), B.H2.as<float>(),
                                                    B.D1.as<float>(), B.D2.as<float>(), B.DY.as<float>(),
                                                    B.L.as<float>(), inv_lab, B.G.as<float>(), B.loss.as<float>());
  EH_HIP(hipGetLastError());
  return EH_OK;
}

}  // namespace

extern "C" eh_status eh_mlp_loss_grad(eh_ctx *ctx, const float *params, const float *x, const float *target,
                                      const float *mask, int batch, float *loss, float *grad) {
  if (!ctx || !params || !x || !target || !mask || !grad || batch <= 0) { eh_set_error("bad argument"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(ctx->device));
  double lab = 0;
  for (int q = 0; q < batch * NO; q++) lab += mask[q];
  const float inv_lab = lab > 0 ? (float)(1.0 / lab) : 0.0f;
  DevBuf P, X, T, Mk;
  TrainBufs B;
  EH_CHECK(B.alloc(batch));
  EH_CHECK(P.alloc(4ull * NPARAM));
  EH_CHECK(X.alloc(4ull * NI * batch));
  EH_CHECK(T.alloc(4ull * NO * batch));
  EH_CHECK(Mk.alloc(4ull * NO * batch));
  EH_HIP(hipMemcpy(P.p, params, 4ull * NPARAM, hipMemcpyHostToDevice));
  EH_HIP(hipMemcpy(X.p, x, 4ull * NI * batch, hipMemcpyHostToDevice));
  EH_HIP(hipMemcpy(T.p, target, 4ull * NO * batch, hipMemcpyHostToDevice));
  EH_HIP(hipMemcpy(Mk.p, mask, 4ull * NO * batch, hipMemcpyHostToDevice));
  EH_CHECK(grad_dev(ctx, P.as<float>(), X.as<float>(), T.as<float>(), Mk.as<float>(), nullptr, batch, inv_lab, B));
  EH_HIP(hipStreamSynchronize(ctx->stream));
  EH_HIP(hipMemcpy(grad, B.G.p, 4ull * NPARAM, hipMemcpyDeviceToHost));
  float l = 0;
  EH_HIP(hipMemcpy(&l, B.loss.p, 4, hipMemcpyDeviceToHost));
  if (loss) *loss = l;
  return EH_OK;
}

extern "C" eh_status eh_selector_train(eh_ctx *ctx, const float *x, const float *target, const float *mask,
                                       int nsamples, int epochs, int batch, float lr, uint64_t seed, float *final_loss) {
  if (!ctx || !x || !target || !mask || nsamples <= 0 || epochs <= 0 || batch <= 0) {
    eh_set_error("bad argument");
    return EH_ERR_INVALID;
  }
  EH_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  batch = std::min(batch, nsamples);
  // He-style initialisation from the seed unless parameters are already set
  std::vector<float> P0(NPARAM, 0.0f);
  if (ctx->mlp.size() == (size_t)NPARAM) {
    P0 = ctx->mlp;
  } else {
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_out, int fan_in) {
      std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (float)fan_in));
      for (int q = 0; q < fan_out * fan_in; q++) P0[off + q] = nd(rng);
    };
    fill(OW1, NH, NI);
    fill(OW2, NH, NH);
    fill(OW3, NO, NH);
  }
  DevBuf P, X, T, Mk, M1, M2, idx;
  TrainBufs B;
  EH_CHECK(B.alloc(batch));
  EH_CHECK(P.alloc(4ull * NPARAM));
  EH_CHECK(M1.alloc(4ull * NPARAM));
  EH_CHECK(M2.alloc(4ull * NPARAM));
  EH_CHECK(X.alloc(4ull * NI * nsamples));
  EH_CHECK(T.alloc(4ull * NO * nsamples));
  EH_CHECK(Mk.alloc(4ull * NO * nsamples));
  EH_CHECK(idx.alloc(4ull * nsamples));
  EH_HIP(hipMemcpy(P.p, P0.data(), 4ull * NPARAM, hipMemcpyHostToDevice));
  EH_HIP(hipMemset(M1.p, 0, 4ull * NPARAM));
  EH_HIP(hipMemset(M2.p, 0, 4ull * NPARAM));
  EH_HIP(hipMemcpy(X.p, x, 4ull * NI * nsamples, hipMemcpyHostToDevice));
  EH_HIP(hipMemcpy(T.p, target, 4ull * NO * nsamples, hipMemcpyHostToDevice));
  EH_HIP(hipMemcpy(Mk.p, mask, 4ull * NO * nsamples, hipMemcpyHostToDevice));
  std::vector<int> perm(nsamples);
  std::iota(perm.begin(), perm.end(), 0);
  std::mt19937_64 rng(seed ^ 0x5EEDull);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  uint64_t step = 0;
  float last = 0;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    EH_HIP(hipMemcpyAsync(idx.p, perm.data(), 4ull * nsamples, hipMemcpyHostToDevice, st));
    for (int b0 = 0; b0 + batch <= nsamples; b0 += batch) {
      double lab = 0;
      for (int q = b0; q < b0 + batch; q++)
        for (int k = 0; k < NO; k++) lab += mask[(size_t)perm[q] * NO + k];
      if (lab <= 0) continue;
      EH_CHECK(grad_dev(ctx, P.as<float>(), X.as<float>(), T.as<float>(), Mk.as<float>(), idx.as<int>() + b0, batch,
                        (float)(1.0 / lab), B));
      step++;
      const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
      k_adam<<<grid_for(NPARAM, 256), 256, 0, st>>>(P.as<float>(), B.G.as<float>(), M1.as<float>(), M2.as<float>(), lr,
                                                     b1, b2, eps, c1, c2);
    }
    EH_HIP(hipMemcpyAsync(&last, B.loss.p, 4, hipMemcpyDeviceToHost, st));
    EH_HIP(hipStreamSynchronize(st));
  }
  EH_HIP(hipGetLastError());
  std::vector<float> out(NPARAM);
  EH_HIP(hipMemcpy(out.data(), P.p, 4ull * NPARAM, hipMemcpyDeviceToHost));
  ctx->mlp = out;
  if (final_loss) *final_loss = last;
  return EH_OK;
}
