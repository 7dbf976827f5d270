// eh_mlp.hip -- the learned per-row selector.
//
// Every row that answers membership queries in the triangle pass can do so with
// one of four arms (RowDesc type values):
//   0 BS  probe one bit of the row's bitset          (intersection.hpp:1023-1058)
//   1 PS  walk the pshort partitions, search the run (intersection.hpp:1062-1173)
//   2 U   binary search in the sorted uint row       (the merge kernels' role)
//   3 UG  galloping search from the front            (set_intersect_galloping, :393-584)
// An FP32 MLP 16 -> 64 -> 64 -> 4 (ReLU) maps 16 row features to the predicted
// log2 shader cycles per query of each arm.  It is trained on timings the engine
// takes itself (k_profile_arms times every arm on the row's real queries), and the
// LEARNED layout policy gives each row the arm with the smallest prediction.  The
// arms only change how a count is computed, never which elements are common, so
// counts are identical under every selection.
//
// Kernels here: feature extraction, arm profiling, MLP forward, MLP loss + per-sample
// deltas, deterministic per-parameter gradient sums, Adam.
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

// ---------------------------------------------------------------- features

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

__device__ __forceinline__ float lg1(float x) { return log2f(1.0f + x); }

// 16 features of row i (all scaled to roughly [0, 1])
__global__ void k_features(const uint32_t *card, const uint32_t *aoff, const uint32_t *unbr, const uint32_t *uses,
                           uint32_t n, float *F) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t c = card[i];
    const uint32_t *r = unbr + (uint64_t)aoff[i] * 4;
    float f[NI];
    for (int q = 0; q < NI; q++) f[q] = 0.0f;
    const float u = (float)uses[i];
    if (c) {
      const uint32_t lo = r[0], hi = r[c - 1];
      const float span = (float)(hi - lo);
      uint32_t parts = 1, low64k = 0, quarter = 0, maxgap = 0;
      const uint32_t qcut = lo + (hi - lo) / 4u;
      for (uint32_t k = 0; k < c; k++) {
        const uint32_t x = r[k];
        if (k) {
          parts += (x >> 16) != (r[k - 1] >> 16);
          maxgap = max(maxgap, x - r[k - 1]);
        }
        low64k += x < 65536u;
        quarter += x <= qcut;
      }
      f[0] = lg1((float)c) / 16.0f;
      f[1] = lg1(span) / 32.0f;
      f[2] = (float)c / (span + 1.0f);
      f[3] = lg1((float)parts) / 16.0f;
      f[6] = lg1((float)((hi >> 6) - (lo >> 6) + 1u)) / 24.0f;
      f[7] = (float)low64k / (float)c;
      f[8] = lg1(span / (float)c) / 32.0f;
      f[9] = lg1((float)maxgap) / 32.0f;
      f[11] = lg1((float)lo) / 32.0f;
      f[12] = lg1((float)hi) / 32.0f;
      f[13] = lg1((float)r[c / 2]) / 32.0f;
      f[14] = (float)quarter / (float)c;
    }
    f[4] = lg1(u) / 24.0f;
    f[5] = lg1(u * (float)c) / 40.0f;
    f[10] = (float)i / (float)n;
    f[15] = 1.0f;
    for (int q = 0; q < NI; q++) F[i * NI + q] = f[q];
  }
}

// ---------------------------------------------------------------- arm profiling

struct ProfArgs {
  const uint32_t *card, *aoff, *unbr;         // the oriented matrix (uint rows)
  const uint64_t *toff;                       // in-neighbour CSR of the oriented matrix
  const uint32_t *tsrc;
  const RowDesc *desc_bs;                     // same rows encoded as bitsets
  const uint64_t *bsw;
  const RowDesc *desc_ps;                     // ... and as pshort streams
  const uint16_t *psd;
  const uint32_t *rows;                       // sampled rows
  uint32_t nrows;
  uint32_t qcap;                              // queries timed per row and arm
  unsigned long long *cycles;                 // nrows x 4
  unsigned long long *queries;                // nrows
  unsigned long long *hits;                   // nrows x 4
};

__device__ __forceinline__ uint32_t arm_member(const ProfArgs &p, uint32_t arm, uint32_t j, uint32_t x) {
  if (arm == EH_BS) {
    const RowDesc d = p.desc_bs[j];
    const uint32_t w = (x >> 6) - d.w0;
    return (w < desc_aux(d)) ? (uint32_t)((p.bsw[d.off + w] >> (x & 63u)) & 1ull) : 0u;
  }
  if (arm == EH_PS) {
    const RowDesc d = p.desc_ps[j];
    const uint16_t *S = p.psd + (uint64_t)d.off * 4;
    const uint32_t len = desc_aux(d), hx = x >> 16;
    uint32_t pos = 0;
    while (pos < len) {
      const uint32_t pre = S[pos], cnt = (uint32_t)S[pos + 1] + 1u;
      if (pre == hx) {
        uint32_t lo = 0, hi = cnt;
        const uint32_t v = x & 0xFFFFu;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (S[pos + 2 + mid] < v) lo = mid + 1; else hi = mid;
        }
        return lo < cnt && S[pos + 2 + lo] == v;
      }
      if (pre > hx) return 0;
      pos += cnt + 2u;
    }
    return 0;
  }
  const uint32_t *R = p.unbr + (uint64_t)p.aoff[j] * 4;
  const uint32_t n = p.card[j];
  if (arm == EH_U) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (R[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo < n && R[lo] == x;
  }
  if (R[0] >= x) return R[0] == x;
  uint32_t b = 1;
  while (b < n && R[b] < x) b <<= 1;
  uint32_t lo = (b >> 1) + 1, hi = min(b, n - 1) + 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (R[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < n && R[lo] == x;
}

// one wave per sampled row j: its queries are the elements below j of every source
// row that contains j; each arm runs twice over the first qcap of them and the
// second run is timed
__global__ void __launch_bounds__(64) k_profile_arms(ProfArgs p) {
  const uint32_t lane = threadIdx.x & 63u;
  for (uint32_t s = blockIdx.x; s < p.nrows; s += gridDim.x) {
    const uint32_t j = p.rows[s];
    const uint64_t t0 = p.toff[j], t1 = p.toff[j + 1];
    uint64_t q_done = 0;
    for (uint32_t arm = 0; arm < 4; arm++) {
      uint64_t cyc = 0, hit = 0;
      for (int rep = 0; rep < 2; rep++) {
        __builtin_amdgcn_s_waitcnt(0);
        const uint64_t c0 = clock64();
        uint32_t h = 0, q = 0;
        for (uint64_t e = t0; e < t1 && q < p.qcap; e++) {
          const uint32_t i = p.tsrc[e];
          const uint32_t d = p.card[i];
          const uint32_t *A = p.unbr + (uint64_t)p.aoff[i] * 4;
          for (uint32_t k = lane; k < d; k += 64) {
            const uint32_t x = A[k];
            if (x < j) h += arm_member(p, arm, j, x);
          }
          uint32_t below = 0;
          for (uint32_t k = lane; k < d; k += 64) below += A[k] < j;
          for (int o = 32; o > 0; o >>= 1) below += __shfl_xor(below, o, 64);
          q += __builtin_amdgcn_readfirstlane(below);
        }
        for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
        const uint64_t c1 = clock64();
        cyc = c1 - c0;
        hit = h;
        q_done = q;
      }
      if (lane == 0) {
        p.cycles[4ull * s + arm] = cyc;
        p.hits[4ull * s + arm] = hit;
      }
    }
    if (lane == 0) p.queries[s] = q_done;
  }
}

// ---------------------------------------------------------------- MLP kernels

__device__ __forceinline__ float relu(float v) { return v > 0.0f ? v : 0.0f; }

// y[b] = MLP(x[b]); one thread per sample
__global__ void k_mlp_forward(const float *__restrict__ P, const float *__restrict__ X, int batch, float *Y) {
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < batch; b += gridDim.x * blockDim.x) {
    float x[NI], h1[NH], y[NO];
    for (int i = 0; i < NI; i++) x[i] = X[(size_t)b * NI + i];
    for (int o = 0; o < NH; o++) {
      float a = P[OB1 + o];
      for (int i = 0; i < NI; i++) a += P[OW1 + o * NI + i] * x[i];
      h1[o] = relu(a);
    }
    for (int k = 0; k < NO; k++) y[k] = P[OB3 + k];
    for (int o = 0; o < NH; o++) {
      float a = P[OB2 + o];
      for (int i = 0; i < NH; i++) a += P[OW2 + o * NH + i] * h1[i];
      const float h2 = relu(a);
      for (int k = 0; k < NO; k++) y[k] += P[OW3 + k * NH + o] * h2;
    }
    for (int k = 0; k < NO; k++) Y[(size_t)b * NO + k] = y[k];
  }
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

// the arm with the smallest predicted cost; bitsets only where the words stay
// within 8 per element (64 bytes, 16x the uint row), rows of one element stay uint
__global__ void k_pick_arm(const float *Y, const uint32_t *card, const uint32_t *aoff, const uint32_t *unbr, uint32_t n,
                           uint8_t *types) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t c = card[i];
    uint8_t t = EH_U;
    if (c > 1) {
      const uint32_t *r = unbr + (uint64_t)aoff[i] * 4;
      const uint64_t nw = (uint64_t)(r[c - 1] >> 6) - (r[0] >> 6) + 1ull;
      float best = 3.0e38f;
      for (uint32_t k = 0; k < NO; k++) {
        if (k == EH_BS && nw > 8ull * c) continue;
        const float v = Y[i * NO + k];
        if (v < best) { best = v; t = (uint8_t)k; }
      }
    }
    types[i] = t;
  }
}

__global__ void k_sample_rows(const uint32_t *card, const uint32_t *uses, uint32_t n, uint32_t stride, uint32_t *flag) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n;
       i += (uint64_t)gridDim.x * blockDim.x)
    flag[i] = (i < n && card[i] >= 2u && uses[i] >= 1u && (i % stride) == 0) ? 1u : 0u;
}

__global__ void k_scatter_flag(const uint32_t *flag, const uint64_t *pos, uint32_t n, uint32_t *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    if (flag[i]) out[pos[i]] = (uint32_t)i;
}

// keep only the sampled rows (compact CSR of the oriented matrix)
__global__ void k_card_if(const uint32_t *card, const uint32_t *flag, uint32_t n, uint32_t *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (i < n && flag[i]) ? card[i] : 0u;
}

__global__ void k_copy_rows(const uint32_t *card, const uint32_t *aoff, const uint32_t *unbr, const uint32_t *flag,
                            const uint64_t *off, uint32_t n, uint32_t *nbr) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = wid; i < n; i += nw) {
    if (flag && !flag[i]) continue;
    const uint32_t c = card[i];
    const uint32_t *r = unbr + (uint64_t)aoff[i] * 4;
    for (uint32_t k = lane; k < c; k += 64) nbr[off[i] + k] = r[k];
  }
}

// transposed oriented CSR: keys (target << 32 | source)
__global__ void k_transpose_keys(const uint32_t *card, const uint32_t *aoff, const uint32_t *unbr, const uint64_t *off,
                                 uint32_t n, uint64_t *keys) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = wid; i < n; i += nw) {
    const uint32_t c = card[i];
    const uint32_t *r = unbr + (uint64_t)aoff[i] * 4;
    for (uint32_t k = lane; k < c; k += 64) keys[off[i] + k] = ((uint64_t)r[k] << 32) | (uint32_t)i;
  }
}

__global__ void k_hi_counts(const uint64_t *keys, uint64_t m, uint32_t *cnt) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[(uint32_t)(keys[e] >> 32)], 1u);
}

__global__ void k_lo_words(const uint64_t *keys, uint64_t m, uint32_t *out) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x)
    out[e] = (uint32_t)keys[e];
}

}  // namespace

// ---------------------------------------------------------------- host helpers

static eh_status features_dev(eh_matrix *m, DevBuf &F, DevBuf &uses) {
  hipStream_t st = m->ctx->stream;
  const uint32_t n = m->n;
  EH_CHECK(uses.alloc(4ull * n + 4));
  EH_HIP(hipMemsetAsync(uses.p, 0, 4ull * n + 4, st));
  k_uses<<<grid_for((uint64_t)n * 64), 256, 0, st>>>(m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(),
                                                     m->dunbr.as<uint32_t>(), n, uses.as<uint32_t>());
  EH_CHECK(F.alloc(4ull * NI * n + 64));
  k_features<<<grid_for(n), 256, 0, st>>>(m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(), m->dunbr.as<uint32_t>(),
                                          uses.as<uint32_t>(), n, F.as<float>());
  EH_HIP(hipGetLastError());
  return EH_OK;
}

static eh_status mlp_forward_dev(eh_ctx *ctx, const float *d_params, const float *d_x, int batch, float *d_y) {
  k_mlp_forward<<<grid_for((uint64_t)batch, 128), 128, 0, ctx->stream>>>(d_params, d_x, batch, d_y);
  EH_HIP(hipGetLastError());
  return EH_OK;
}

eh_status matrix_select_learned(eh_matrix *m, uint8_t *d_types) {
  eh_ctx *ctx = m->ctx;
  if (ctx->mlp.size() != (size_t)NPARAM) {
    eh_set_error("EH_LAYOUT_LEARNED needs trained selector parameters (eh_selector_set_params / eh_selector_train)");
    return EH_ERR_INVALID;
  }
  DevBuf F, uses, P, Y;
  EH_CHECK(features_dev(m, F, uses));
  EH_CHECK(P.alloc(4ull * NPARAM));
  EH_HIP(hipMemcpy(P.p, ctx->mlp.data(), 4ull * NPARAM, hipMemcpyHostToDevice));
  EH_CHECK(Y.alloc(4ull * NO * m->n + 64));
  EH_CHECK(mlp_forward_dev(ctx, P.as<float>(), F.as<float>(), (int)m->n, Y.as<float>()));
  k_pick_arm<<<grid_for(m->n), 256, 0, ctx->stream>>>(Y.as<float>(), m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(),
                                                      m->dunbr.as<uint32_t>(), m->n, d_types);
  EH_HIP(hipGetLastError());
  EH_HIP(hipStreamSynchronize(ctx->stream));
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

extern "C" eh_status eh_matrix_features(eh_matrix *m, float *features) {
  if (!m || !features) { eh_set_error("null argument"); return EH_ERR_INVALID; }
  EH_HIP(hipSetDevice(m->ctx->device));
  DevBuf F, uses;
  EH_CHECK(features_dev(m, F, uses));
  EH_HIP(hipStreamSynchronize(m->ctx->stream));
  EH_HIP(hipMemcpy(features, F.p, 4ull * NI * m->n, hipMemcpyDeviceToHost));
  return EH_OK;
}

extern "C" eh_status eh_matrix_profile_layouts(eh_matrix *m, uint32_t stride, float *targets, float *mask) {
  if (!m || !targets || !mask || stride == 0) { eh_set_error("bad argument"); return EH_ERR_INVALID; }
  if (m->filter != EH_FILTER_LOWER) { eh_set_error("arm profiling needs an oriented (EH_FILTER_LOWER) matrix"); return EH_ERR_INVALID; }
  eh_ctx *ctx = m->ctx;
  EH_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const uint32_t n = m->n;
  DevBuf F, uses, flag, fpos, rows;
  EH_CHECK(features_dev(m, F, uses));
  EH_CHECK(flag.alloc(4ull * n + 8));
  EH_CHECK(fpos.alloc(8ull * n + 16));
  k_sample_rows<<<grid_for((uint64_t)n + 1), 256, 0, st>>>(m->dcard.as<uint32_t>(), uses.as<uint32_t>(), n, stride,
                                                          flag.as<uint32_t>());
  EH_CHECK(prim_exscan_u32_to_u64(ctx, flag.as<uint32_t>(), fpos.as<uint64_t>(), (uint64_t)n + 1));
  uint64_t ns = 0;
  EH_HIP(hipMemcpy(&ns, fpos.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost));
  for (uint64_t q = 0; q < 4ull * n; q++) { targets[q] = 0.0f; mask[q] = 0.0f; }
  if (ns == 0) return EH_OK;
  EH_CHECK(rows.alloc(4 * ns + 16));
  k_scatter_flag<<<grid_for(n), 256, 0, st>>>(flag.as<uint32_t>(), fpos.as<uint64_t>(), n, rows.as<uint32_t>());
  // compact CSR of the sampled rows, encoded once as bitsets and once as pshort
  DevBuf cs, offs, nbrs;
  EH_CHECK(cs.alloc(4ull * n + 8));
  EH_CHECK(offs.alloc(8ull * n + 16));
  k_card_if<<<grid_for((uint64_t)n + 1), 256, 0, st>>>(m->dcard.as<uint32_t>(), flag.as<uint32_t>(), n, cs.as<uint32_t>());
  EH_CHECK(prim_exscan_u32_to_u64(ctx, cs.as<uint32_t>(), offs.as<uint64_t>(), (uint64_t)n + 1));
  uint64_t es = 0;
  EH_HIP(hipMemcpy(&es, offs.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost));
  EH_CHECK(nbrs.alloc(4 * es + 16));
  k_copy_rows<<<grid_for((uint64_t)n * 64), 256, 0, st>>>(m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(),
                                                          m->dunbr.as<uint32_t>(), flag.as<uint32_t>(),
                                                          offs.as<uint64_t>(), n, nbrs.as<uint32_t>());
  eh_matrix *Mb = nullptr, *Mp = nullptr;
  EH_CHECK(matrix_build_dev(ctx, n, offs.as<uint64_t>(), nbrs.as<uint32_t>(), es, EH_FILTER_ALL, nullptr,
                            EH_LAYOUT_BITSET, nullptr, &Mb));
  eh_status s = matrix_build_dev(ctx, n, offs.as<uint64_t>(), nbrs.as<uint32_t>(), es, EH_FILTER_ALL, nullptr,
                                 EH_LAYOUT_PSHORT, nullptr, &Mp);
  if (s != EH_OK) { eh_matrix_free(Mb); return s; }
  // in-neighbours of every row in the oriented matrix
  DevBuf coff, keys, keys2, tcnt, toff, tsrc;
  uint64_t mm = m->card;
  s = coff.alloc(8ull * n + 16);
  if (s == EH_OK) s = keys.alloc(8 * mm + 16);
  if (s == EH_OK) s = keys2.alloc(8 * mm + 16);
  if (s == EH_OK) s = tcnt.alloc(4ull * n + 8);
  if (s == EH_OK) s = toff.alloc(8ull * n + 16);
  if (s == EH_OK) s = tsrc.alloc(4 * mm + 16);
  if (s == EH_OK) s = prim_exscan_u32_to_u64(ctx, m->dcard.as<uint32_t>(), coff.as<uint64_t>(), (uint64_t)n);
  if (s == EH_OK) {
    k_transpose_keys<<<grid_for((uint64_t)n * 64), 256, 0, st>>>(m->dcard.as<uint32_t>(), m->daoff.as<uint32_t>(),
                                                                 m->dunbr.as<uint32_t>(), coff.as<uint64_t>(), n,
                                                                 keys.as<uint64_t>());
    s = prim_sort_u64(ctx, keys.as<uint64_t>(), keys2.as<uint64_t>(), mm, 64);
  }
  if (s == EH_OK) {
    (void)hipMemsetAsync(tcnt.p, 0, 4ull * n + 8, st);
    k_hi_counts<<<grid_for(mm), 256, 0, st>>>(keys.as<uint64_t>(), mm, tcnt.as<uint32_t>());
    s = prim_exscan_u32_to_u64(ctx, tcnt.as<uint32_t>(), toff.as<uint64_t>(), (uint64_t)n + 1);
  }
  DevBuf cyc, qs, hits;
  if (s == EH_OK) {
    k_lo_words<<<grid_for(mm), 256, 0, st>>>(keys.as<uint64_t>(), mm, tsrc.as<uint32_t>());
    s = cyc.alloc(32 * ns + 64);
    if (s == EH_OK) s = qs.alloc(8 * ns + 64);
    if (s == EH_OK) s = hits.alloc(32 * ns + 64);
  }
  if (s == EH_OK) {
    ProfArgs p;
    p.card = m->dcard.as<uint32_t>();
    p.aoff = m->daoff.as<uint32_t>();
    p.unbr = m->dunbr.as<uint32_t>();
    p.toff = toff.as<uint64_t>();
    p.tsrc = tsrc.as<uint32_t>();
    p.desc_bs = Mb->ddesc.as<RowDesc>();
    p.bsw = Mb->dbsw.as<uint64_t>();
    p.desc_ps = Mp->ddesc.as<RowDesc>();
    p.psd = Mp->dpsd.as<uint16_t>();
    p.rows = rows.as<uint32_t>();
    p.nrows = (uint32_t)ns;
    p.qcap = 2048;
    p.cycles = cyc.as<unsigned long long>();
    p.queries = qs.as<unsigned long long>();
    p.hits = hits.as<unsigned long long>();
    const unsigned blocks = (unsigned)std::min<uint64_t>(ns, (uint64_t)(ctx->num_cus ? ctx->num_cus : 256) * 4);
    k_profile_arms<<<blocks, 64, 0, st>>>(p);
    if (hipGetLastError() != hipSuccess) s = EH_ERR_HIP;
  }
  std::vector<uint32_t> hr(ns);
  std::vector<uint64_t> hc(4 * ns), hq(ns), hh(4 * ns);
  if (s == EH_OK && (hipMemcpy(hr.data(), rows.p, 4 * ns, hipMemcpyDeviceToHost) != hipSuccess ||
                     hipMemcpy(hc.data(), cyc.p, 32 * ns, hipMemcpyDeviceToHost) != hipSuccess ||
                     hipMemcpy(hq.data(), qs.p, 8 * ns, hipMemcpyDeviceToHost) != hipSuccess ||
                     hipMemcpy(hh.data(), hits.p, 32 * ns, hipMemcpyDeviceToHost) != hipSuccess))
    s = EH_ERR_HIP;
  eh_matrix_free(Mb);
  eh_matrix_free(Mp);
  if (s != EH_OK) return s;
  for (uint64_t q = 0; q < ns; q++) {
    // every arm must find the same common elements
    if (hh[4 * q] != hh[4 * q + 1] || hh[4 * q] != hh[4 * q + 2] || hh[4 * q] != hh[4 * q + 3]) {
      eh_set_error("arm profiling: arms disagree on row %u", hr[q]);
      return EH_ERR_HIP;
    }
    if (hq[q] < 8) continue;
    const uint32_t r = hr[q];
    for (int k = 0; k < 4; k++) {
      targets[4ull * r + k] = (float)std::log2((double)hc[4 * q + k] / (double)hq[q]);
      mask[4ull * r + k] = 1.0f;
    }
  }
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
