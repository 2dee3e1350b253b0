// mlp.hip — the learned schedule's 16-64-64-2 ReLU MLP: forward, backward
// and SGD kernels (bf16 weights, fp32 accumulation), plus the per-level
// decision kernel that turns the device-resident level counters and the
// frontier's per-page activity into the model's features and runs the
// forward pass, so the direction for the next level comes back to the host
// with the level counters in one copy.  No reference counterpart.
//
// Parameter layout (bf16 bit patterns or fp32 master copy):
//   W1[64][16] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]   (5378 values)
#include "common.h"

namespace {

constexpr int kIn = 16, kH = 64, kOut = 2;
constexpr int kOffB1 = kH * kIn;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kOut * kH;
constexpr int kNParam = kOffB3 + kOut;
static_assert(kNParam == SCAPH_MLP_NPARAM, "parameter count");
constexpr uint32_t kBlock = 256;

// bf16 is the top half of an IEEE binary32.
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}

// Round-to-nearest-even onto the bf16 grid; NaNs stay (quiet) NaNs.
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ void stage_params(float* w, const uint16_t* p) {
  for (int i = threadIdx.x; i < kNParam; i += blockDim.x) w[i] = bf16_to_f32(p[i]);
  __syncthreads();
}

// One sample on one wave: lane j owns hidden unit j.  h1/h2 are written
// for the backward pass when not null.  Returns the outputs in y[0..1]
// (valid on every lane).
__device__ __forceinline__ void forward_wave(const float* w, const float* x, float* h1_out,
                                             float* h2_out, float* y) {
  const int j = (int)lane_id();
  float a = w[kOffB1 + j];
#pragma unroll
  for (int i = 0; i < kIn; ++i) a += w[j * kIn + i] * x[i];
  const float h1 = a > 0.f ? a : 0.f;
  float b = w[kOffB2 + j];
  for (int i = 0; i < kH; ++i) b += w[kOffW2 + j * kH + i] * __shfl(h1, i, SCAPH_WAVE);
  const float h2 = b > 0.f ? b : 0.f;
  if (h1_out) h1_out[j] = h1;
  if (h2_out) h2_out[j] = h2;
#pragma unroll
  for (int o = 0; o < kOut; ++o) y[o] = wave_sum(w[kOffW3 + o * kH + j] * h2) + w[kOffB3 + o];
}

// The same forward pass reading the bf16 parameters straight from global
// memory: lane j issues all loads of its W1 and W2 rows (16 B vectors) at
// once, so a single decision costs one memory latency instead of staging
// all 5378 parameters through LDS.
__device__ __forceinline__ void unpack8(const uint4 q, float* f) {
  const uint32_t u[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    f[2 * k] = __uint_as_float(u[k] << 16);
    f[2 * k + 1] = __uint_as_float(u[k] & 0xFFFF0000u);
  }
}

struct RowRegs {
  uint4 q1[2], q2[8];
  float b1, b2, w30, w31, b30, b31;
};

// Lane j's W1 and W2 rows and the scalars it needs, issued together.
__device__ __forceinline__ void load_rows(const uint16_t* p, RowRegs& r) {
  const int j = (int)lane_id();
  const uint4* r1 = reinterpret_cast<const uint4*>(p + j * kIn);
  const uint4* r2 = reinterpret_cast<const uint4*>(p + kOffW2 + j * kH);
#pragma unroll
  for (int k = 0; k < 2; ++k) r.q1[k] = r1[k];
#pragma unroll
  for (int k = 0; k < 8; ++k) r.q2[k] = r2[k];
  r.b1 = bf16_to_f32(p[kOffB1 + j]);
  r.b2 = bf16_to_f32(p[kOffB2 + j]);
  r.w30 = bf16_to_f32(p[kOffW3 + j]);
  r.w31 = bf16_to_f32(p[kOffW3 + kH + j]);
  r.b30 = bf16_to_f32(p[kOffB3]);
  r.b31 = bf16_to_f32(p[kOffB3 + 1]);
}

// forward_wave's arithmetic (same order of accumulation) on loaded rows;
// unit i's h1 reaches every lane by v_readlane (a scalar operand of the
// FMA) instead of an LDS permute per unit.
__device__ __forceinline__ void forward_rows(const RowRegs& r, const float* x, float* y) {
  float w1[kIn];
  unpack8(r.q1[0], w1);
  unpack8(r.q1[1], w1 + 8);
  float a = r.b1;
#pragma unroll
  for (int i = 0; i < kIn; ++i) a += w1[i] * x[i];
  const float h1 = a > 0.f ? a : 0.f;
  float b = r.b2;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    float w2[8];
    unpack8(r.q2[k], w2);
#pragma unroll
    for (int i = 0; i < 8; ++i)
      b += w2[i] * __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(h1), 8 * k + i));
  }
  const float h2 = b > 0.f ? b : 0.f;
  y[0] = wave_sum(r.w30 * h2) + r.b30;
  y[1] = wave_sum(r.w31 * h2) + r.b31;
}

__global__ void __launch_bounds__(kBlock)
k_mlp_forward(const uint16_t* params, const float* x, uint32_t batch, float* y, float* act) {
  __shared__ float w[kNParam];
  stage_params(w, params);
  const uint32_t wid = threadIdx.x / SCAPH_WAVE;
  const uint32_t nw = blockDim.x / SCAPH_WAVE;
  for (uint32_t s = blockIdx.x * nw + wid; s < batch; s += gridDim.x * nw) {
    float xs[kIn];
#pragma unroll
    for (int i = 0; i < kIn; ++i) xs[i] = x[(uint64_t)s * kIn + i];
    float out[kOut];
    forward_wave(w, xs, act ? act + (uint64_t)s * 2 * kH : nullptr,
                 act ? act + (uint64_t)s * 2 * kH + kH : nullptr, out);
    if (lane_id() == 0) {
      y[(uint64_t)s * kOut] = out[0];
      y[(uint64_t)s * kOut + 1] = out[1];
    }
  }
}

__global__ void k_zero_f32(float* p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = 0.f;
}

// Batch-mean gradient of  softplus(z0) - t0*z0 + w1*(z1 - t1)^2.
// Each block accumulates its samples' gradients in LDS, then adds them to
// the global gradient (zeroed beforehand) with one atomic per parameter.
__global__ void __launch_bounds__(kBlock)
k_mlp_backward(const uint16_t* params, const float* x, const float* act, const float* y,
               const float* label, uint32_t batch, float w1, float* grad, float* loss) {
  __shared__ float w[kNParam];
  __shared__ float g[kNParam];
  for (int i = threadIdx.x; i < kNParam; i += blockDim.x) g[i] = 0.f;
  stage_params(w, params);
  const uint32_t wid = threadIdx.x / SCAPH_WAVE;
  const uint32_t nw = blockDim.x / SCAPH_WAVE;
  const int j = (int)lane_id();
  const float inv_b = 1.f / (float)batch;
  float lsum = 0.f;
  for (uint32_t s = blockIdx.x * nw + wid; s < batch; s += gridDim.x * nw) {
    const float* h1p = act + (uint64_t)s * 2 * kH;
    const float h1 = h1p[j], h2 = h1p[kH + j];
    const float z0 = y[(uint64_t)s * kOut], z1 = y[(uint64_t)s * kOut + 1];
    const float t0 = label[(uint64_t)s * kOut], t1 = label[(uint64_t)s * kOut + 1];
    const float sig = 1.f / (1.f + __expf(-z0));
    const float dz0 = (sig - t0) * inv_b;
    const float dz1 = 2.f * w1 * (z1 - t1) * inv_b;
    if (j == 0) {
      const float sp = z0 > 0.f ? z0 + __logf(1.f + __expf(-z0)) : __logf(1.f + __expf(z0));
      lsum += sp - t0 * z0 + w1 * (z1 - t1) * (z1 - t1);
      atomicAdd(&g[kOffB3], dz0);
      atomicAdd(&g[kOffB3 + 1], dz1);
    }
    // layer 3
    atomicAdd(&g[kOffW3 + j], dz0 * h2);
    atomicAdd(&g[kOffW3 + kH + j], dz1 * h2);
    float d2 = w[kOffW3 + j] * dz0 + w[kOffW3 + kH + j] * dz1;
    d2 = h2 > 0.f ? d2 : 0.f;
    // layer 2: unit j's row of W2 times h1 (from the lanes)
    atomicAdd(&g[kOffB2 + j], d2);
    for (int i = 0; i < kH; ++i) atomicAdd(&g[kOffW2 + j * kH + i], d2 * __shfl(h1, i, SCAPH_WAVE));
    // back to h1: d1_j = sum_k W2[k][j] d2_k
    float d1 = 0.f;
    for (int k = 0; k < kH; ++k) d1 += w[kOffW2 + k * kH + j] * __shfl(d2, k, SCAPH_WAVE);
    d1 = h1 > 0.f ? d1 : 0.f;
    atomicAdd(&g[kOffB1 + j], d1);
    const float* xs = x + (uint64_t)s * kIn;
#pragma unroll
    for (int i = 0; i < kIn; ++i) atomicAdd(&g[j * kIn + i], d1 * xs[i]);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kNParam; i += blockDim.x)
    if (g[i] != 0.f) atomicAdd(&grad[i], g[i]);
  if (j == 0 && lsum != 0.f) atomicAdd(loss, lsum * inv_b);
}

__global__ void __launch_bounds__(kBlock)
k_mlp_sgd(float* master, float* mom, const float* grad, uint32_t n, float lr, float mu,
          uint16_t* params) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float v = mu * mom[i] + grad[i];
    mom[i] = v;
    const float m = master[i] - lr * v;
    master[i] = m;
    params[i] = f32_to_bf16(m);
  }
}

// ---------------------------------------------------------------- decision

// Per-page active-vertex counts of a frontier given as a bitmap or a list,
// then (last block to finish) the 16 features, the forward pass and the
// decision, written into `out`:
//   out->logit_pull, out->threshold, out->pages_active, out->hist[12]
// Features (same as the host's level_features): log2(1+n_f), log2(1+m_f),
// log2(1+m_u), level, then the 12-bin histogram of the per-page active
// share (bin 0: none; bins 1..11: floor(log2(share)) + 11 clamped to 0..10,
// plus one), normalised by (x - mean) / scale.
struct DecideOut {
  float logit_pull;
  float threshold;
  float pages_active;
  float pad;
  float hist[12];
  float feat[16];
};

__global__ void __launch_bounds__(kBlock)
k_frontier_pages(const uint32_t* bitmap, const uint32_t* queue,
                 const unsigned long long* queue_len, const scaph_page* pages, uint32_t num_pages,
                 uint32_t n, uint32_t* page_cnt) {
  // one block per page when the frontier is a bitmap; grid-stride over the
  // list otherwise
  if (bitmap) {
    __shared__ uint32_t red[kBlock / SCAPH_WAVE];
    for (uint32_t p = blockIdx.x; p < num_pages; p += gridDim.x) {
      const uint32_t left = pages[p].left, right = pages[p].right;
      const uint32_t w0 = left >> 5, w1 = right >> 5;
      uint32_t c = 0;
      for (uint32_t w = w0 + threadIdx.x; w <= w1; w += kBlock) {
        uint32_t bits = bitmap[w];
        if (w == w0) bits &= ~0u << (left & 31);
        if (w == w1) bits &= (right & 31) == 31 ? ~0u : ((1u << ((right & 31) + 1)) - 1u);
        c += __popc(bits);
      }
      c = wave_sum(c);
      if (lane_id() == 0) red[threadIdx.x / SCAPH_WAVE] = c;
      __syncthreads();
      if (threadIdx.x == 0) page_cnt[p] = red[0] + red[1] + red[2] + red[3];
      __syncthreads();
    }
  } else {
    const unsigned long long n_queue = *queue_len;
    for (unsigned long long i = blockIdx.x * blockDim.x + threadIdx.x; i < n_queue;
         i += (unsigned long long)gridDim.x * blockDim.x) {
      const uint32_t v = queue[i];
      if (v >= n) continue;
      uint32_t lo = 0, hi = num_pages;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pages[mid].left <= v) lo = mid; else hi = mid;
      }
      atomicAdd(&page_cnt[lo], 1u);
    }
  }
}

// List form with at most kLdsPages pages: page bounds and a block-private
// histogram in LDS, so the global counters see one add per (block, page)
// instead of one per vertex (many vertices of a frontier share a page).
constexpr uint32_t kLdsPages = 4096;

__global__ void __launch_bounds__(kBlock)
k_queue_pages(const uint32_t* queue, const unsigned long long* queue_len,
              const scaph_page* pages, uint32_t num_pages, uint32_t n, uint32_t* page_cnt) {
  __shared__ uint32_t left[kLdsPages];
  __shared__ uint32_t hist[kLdsPages];
  for (uint32_t p = threadIdx.x; p < num_pages; p += blockDim.x) {
    left[p] = pages[p].left;
    hist[p] = 0;
  }
  __syncthreads();
  const unsigned long long n_queue = *queue_len;
  for (unsigned long long i = blockIdx.x * blockDim.x + threadIdx.x; i < n_queue;
       i += (unsigned long long)gridDim.x * blockDim.x) {
    const uint32_t v = queue[i];
    if (v >= n) continue;
    uint32_t lo = 0, hi = num_pages;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (left[mid] <= v) lo = mid; else hi = mid;
    }
    atomicAdd(&hist[lo], 1u);
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < num_pages; p += blockDim.x)
    if (hist[p]) atomicAdd(&page_cnt[p], hist[p]);
}

// The decision proper, by one block: per-page shares -> 12-bin histogram,
// the 16 features, normalisation, forward pass.  `w` holds the staged
// parameters; page_cnt must be visible to this block.
__device__ void decide_block(const uint16_t* params, const float* norm, const uint32_t* page_cnt,
                             const scaph_page* pages, uint32_t num_pages,
                             const scaph_level_counters* cnt, unsigned long long m_u,
                             uint32_t level, DecideOut* out, float* hist, float* active) {
  // everything the forward pass needs is requested before the page loop
  RowRegs rows;
  float nrm[2 * kIn];
  unsigned long long n_next = 0, m_next = 0;
  if (threadIdx.x < SCAPH_WAVE) {
    load_rows(params, rows);
    const float4* n4 = reinterpret_cast<const float4*>(norm);
#pragma unroll
    for (int k = 0; k < 2 * kIn / 4; ++k) {
      const float4 t = n4[k];
      nrm[4 * k] = t.x;
      nrm[4 * k + 1] = t.y;
      nrm[4 * k + 2] = t.z;
      nrm[4 * k + 3] = t.w;
    }
    n_next = cnt->n_next;
    m_next = cnt->m_next;
  }
  if (threadIdx.x < 12) hist[threadIdx.x] = 0.f;
  if (threadIdx.x == 0) *active = 0.f;
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < num_pages; p += blockDim.x) {
    const float share = (float)page_cnt[p] / (float)(pages[p].nodenum ? pages[p].nodenum : 1u);
    int bin = 0;
    if (share > 0.f) {
      const float l = floorf(__log2f(share) + 11.f);
      bin = 1 + (int)fminf(10.f, fmaxf(0.f, l));
      atomicAdd(active, 1.f);
    }
    atomicAdd(&hist[bin], 1.f);
  }
  __syncthreads();
  if (threadIdx.x < SCAPH_WAVE) {
    const float np = (float)(num_pages ? num_pages : 1u);
    float x[kIn];
    x[0] = __log2f(1.f + (float)n_next);
    x[1] = __log2f(1.f + (float)m_next);
    const unsigned long long mu = m_u > m_next ? m_u - m_next : 0ull;
    x[2] = __log2f(1.f + (float)mu);
    x[3] = (float)level;
    for (int i = 0; i < 12; ++i) x[4 + i] = hist[i] / np;
    float xn[kIn];
#pragma unroll
    for (int i = 0; i < kIn; ++i) {
      const float sc = nrm[kIn + i];
      xn[i] = (x[i] - nrm[i]) / (sc != 0.f ? sc : 1.f);
    }
    float y[kOut];
    forward_rows(rows, xn, y);
    if (threadIdx.x == 0) {
      out->logit_pull = y[0];
      out->threshold = y[1];
      out->pages_active = *active;
      out->pad = 0.f;
      for (int i = 0; i < 12; ++i) out->hist[i] = hist[i] / np;
      for (int i = 0; i < kIn; ++i) out->feat[i] = x[i];
    }
  }
}

// The decision of a model that reads only the four scalar features (its
// W1 columns for the histogram are zero), fused with the level hand-off:
// one wave forms the features from the counters, runs the forward pass and
// copies `words` words of src to the host-mapped dst with the DecideOut
// record at word decide_word, clears the first clear_words of src, then
// releases the flag as scaph_publish_words does.  No frontier pass and no
// extra launch per level.
__global__ void __launch_bounds__(SCAPH_WAVE)
k_decide_publish(const uint16_t* params, const float* norm, uint32_t* src, uint32_t decide_word,
                 unsigned long long m_u, uint32_t level, uint32_t words, uint32_t clear_words,
                 uint32_t* dst, uint32_t* flag, uint32_t seq) {
  __shared__ float rec[sizeof(DecideOut) / 4];
  RowRegs rows;
  load_rows(params, rows);
  const float4* n4 = reinterpret_cast<const float4*>(norm);
  float nrm[2 * kIn];
#pragma unroll
  for (int k = 0; k < 2 * kIn / 4; ++k) {
    const float4 t = n4[k];
    nrm[4 * k] = t.x;
    nrm[4 * k + 1] = t.y;
    nrm[4 * k + 2] = t.z;
    nrm[4 * k + 3] = t.w;
  }
  const scaph_level_counters* cnt = reinterpret_cast<const scaph_level_counters*>(src);
  const unsigned long long n_next = cnt->n_next, m_next = cnt->m_next;
  float x[kIn];
  x[0] = __log2f(1.f + (float)n_next);
  x[1] = __log2f(1.f + (float)m_next);
  x[2] = __log2f(1.f + (float)(m_u > m_next ? m_u - m_next : 0ull));
  x[3] = (float)level;
#pragma unroll
  for (int i = 4; i < kIn; ++i) x[i] = 0.f;
  float xn[kIn];
#pragma unroll
  for (int i = 0; i < kIn; ++i) {
    const float sc = nrm[kIn + i];
    xn[i] = (x[i] - nrm[i]) / (sc != 0.f ? sc : 1.f);
  }
  float y[kOut];
  forward_rows(rows, xn, y);
  const uint32_t lane = threadIdx.x;
  constexpr uint32_t kRec = sizeof(DecideOut) / 4;
  if (lane < kRec) {
    float v = 0.f;
    if (lane == 0) v = y[0];
    else if (lane == 1) v = y[1];
    else if (lane >= 16 && lane < 16 + kIn) v = x[lane - 16];
    rec[lane] = v;
  }
  __syncthreads();
  for (uint32_t i = lane; i < words; i += SCAPH_WAVE) {
    const bool in_rec = i >= decide_word && i < decide_word + kRec;
    dst[i] = in_rec ? __float_as_uint(rec[i - decide_word]) : src[i];
    if (i < clear_words) src[i] = 0u;
  }
  __threadfence_system();
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void __launch_bounds__(kBlock)
k_decide(const uint16_t* params, const float* norm /* mean[16], scale[16] */,
         const uint32_t* page_cnt, const scaph_page* pages, uint32_t num_pages,
         const scaph_level_counters* cnt, unsigned long long m_u, uint32_t level,
         DecideOut* out) {
  __shared__ float hist[12];
  __shared__ float active;
  decide_block(params, norm, page_cnt, pages, num_pages, cnt, m_u, level, out, hist, &active);
}

// Per-page frontier counts added into page_cnt (zero on entry): every block
// counts its share of the frontier (bitmap: 64 vertices per thread, split
// at page boundaries; list: one vertex per thread) into an LDS histogram
// and flushes it with one atomic per touched page.
__global__ void __launch_bounds__(kBlock)
k_count_frontier(const uint32_t* bitmap, const uint32_t* queue, const scaph_page* pages,
                 uint32_t num_pages, uint32_t n, const scaph_level_counters* cnt,
                 uint32_t* page_cnt) {
  __shared__ uint32_t left[kLdsPages];
  __shared__ uint32_t lhist[kLdsPages];
  const bool lds = num_pages <= kLdsPages;
  if (lds) {
    for (uint32_t p = threadIdx.x; p < num_pages; p += blockDim.x) {
      left[p] = pages[p].left;
      lhist[p] = 0;
    }
  }
  __syncthreads();
  auto page_of = [&](uint32_t v) {
    uint32_t lo = 0, hi = num_pages;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if ((lds ? left[mid] : pages[mid].left) <= v) lo = mid; else hi = mid;
    }
    return lo;
  };
  auto add = [&](uint32_t p, uint32_t c) {
    if (!c) return;
    if (lds) atomicAdd(&lhist[p], c);
    else atomicAdd(&page_cnt[p], c);
  };
  if (bitmap) {
    // one 64-vertex pair of words per thread; a pair that straddles a page
    // boundary is split at the boundary
    const uint32_t pairs = (n + 63) >> 6;
    const uint2* bm2 = reinterpret_cast<const uint2*>(bitmap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < pairs;
         i += gridDim.x * blockDim.x) {
      const uint2 q = bm2[i];
      uint64_t bits = ((uint64_t)q.y << 32) | q.x;
      const uint32_t v0 = i << 6;
      if (n - v0 < 64) bits &= (1ull << (n - v0)) - 1ull;
      if (!bits) continue;
      uint32_t p = page_of(v0);
      while (bits) {
        const uint32_t right = pages[p].right;  // inclusive
        const uint32_t span = right - v0 + 1;  // vertices of this pair in page p and before
        const uint64_t in_p = span >= 64 ? bits : bits & ((1ull << span) - 1ull);
        add(p, (uint32_t)__popcll(in_p));
        bits &= ~in_p;
        if (++p >= num_pages) break;
      }
    }
  } else {
    const unsigned long long n_queue = cnt->q_len;
    for (unsigned long long i = blockIdx.x * blockDim.x + threadIdx.x; i < n_queue;
         i += (unsigned long long)gridDim.x * blockDim.x) {
      const uint32_t v = queue[i];
      if (v >= n) continue;
      add(page_of(v), 1u);
    }
  }
  __syncthreads();
  if (lds)
    for (uint32_t p = threadIdx.x; p < num_pages; p += blockDim.x)
      if (lhist[p]) atomicAdd(&page_cnt[p], lhist[p]);
}

// The decision after k_count_frontier: one block; it clears page_cnt
// behind itself so the next level's counts start from zero.
__global__ void __launch_bounds__(kBlock)
k_decide_clear(const uint16_t* params, const float* norm, uint32_t* page_cnt,
               const scaph_page* pages, uint32_t num_pages, const scaph_level_counters* cnt,
               unsigned long long m_u, uint32_t level, DecideOut* out) {
  __shared__ float hist[12];
  __shared__ float active;
  decide_block(params, norm, page_cnt, pages, num_pages, cnt, m_u, level, out, hist, &active);
  for (uint32_t p = threadIdx.x; p < num_pages; p += blockDim.x) page_cnt[p] = 0;
}

inline unsigned grid_for(size_t items, size_t per_block, unsigned cap) {
  size_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

}  // namespace

extern "C" {

int scaph_mlp_forward(const uint16_t* params_bf16, const float* x, uint32_t batch, float* y,
                      float* act, void* stream) {
  if (!params_bf16 || !x || !y) return SCAPH_STATUS_INVALID;
  if (batch == 0) return 0;
  hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for(batch, 4, 1024)), dim3(kBlock), 0,
                     as_stream(stream), params_bf16, x, batch, y, act);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_mlp_backward(const uint16_t* params_bf16, const float* x, const float* act,
                       const float* y, const float* label, uint32_t batch, float w1, float* grad,
                       float* loss, void* stream) {
  if (!params_bf16 || !x || !act || !y || !label || !grad || !loss) return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_zero_f32, dim3(22), dim3(kBlock), 0, st, grad, (uint32_t)kNParam);
  SCAPH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, st, loss, 1u);
  SCAPH_LAUNCH_CHECK();
  if (batch == 0) return 0;
  hipLaunchKernelGGL(k_mlp_backward, dim3(grid_for(batch, 16, 64)), dim3(kBlock), 0, st,
                     params_bf16, x, act, y, label, batch, w1, grad, loss);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_mlp_sgd(float* master, float* momentum, const float* grad, uint32_t nparam, float lr,
                  float mu, uint16_t* params_bf16, void* stream) {
  if (!master || !momentum || !grad || !params_bf16) return SCAPH_STATUS_INVALID;
  hipLaunchKernelGGL(k_mlp_sgd, dim3(grid_for(nparam, kBlock, 64)), dim3(kBlock), 0,
                     as_stream(stream), master, momentum, grad, nparam, lr, mu, params_bf16);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_frontier_page_counts(const uint32_t* bitmap, const uint32_t* queue,
                               const unsigned long long* queue_len, const scaph_page* pages,
                               uint32_t num_pages, uint32_t n, uint32_t* page_cnt, void* stream) {
  if (!pages || !page_cnt || (!bitmap && (!queue || !queue_len))) return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  if (!bitmap) {
    hipLaunchKernelGGL(k_zero_f32, dim3(grid_for(num_pages, kBlock, 64)), dim3(kBlock), 0, st,
                       reinterpret_cast<float*>(page_cnt), num_pages);
    SCAPH_LAUNCH_CHECK();
    if (num_pages <= kLdsPages)
      hipLaunchKernelGGL(k_queue_pages, dim3(256), dim3(kBlock), 0, st, queue, queue_len, pages,
                         num_pages, n, page_cnt);
    else
      hipLaunchKernelGGL(k_frontier_pages, dim3(1024), dim3(kBlock), 0, st, bitmap, queue,
                         queue_len, pages, num_pages, n, page_cnt);
  } else {
    hipLaunchKernelGGL(k_frontier_pages, dim3(grid_for(num_pages, 1, 4096)), dim3(kBlock), 0, st,
                       bitmap, queue, queue_len, pages, num_pages, n, page_cnt);
  }
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_mlp_decide(const uint16_t* params_bf16, const float* norm, const uint32_t* page_cnt,
                     const scaph_page* pages, uint32_t num_pages,
                     const scaph_level_counters* counters, unsigned long long m_u,
                     uint32_t level, float* out, void* stream) {
  if (!params_bf16 || !norm || !page_cnt || !pages || !counters || !out)
    return SCAPH_STATUS_INVALID;
  hipLaunchKernelGGL(k_decide, dim3(1), dim3(kBlock), 0, as_stream(stream), params_bf16, norm,
                     page_cnt, pages, num_pages, counters, m_u, level,
                     reinterpret_cast<DecideOut*>(out));
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_mlp_decide_publish(const uint16_t* params_bf16, const float* norm, uint32_t* counters,
                             uint32_t decide_word, unsigned long long m_u, uint32_t level,
                             uint32_t words, uint32_t clear_words, uint32_t* dst_mapped,
                             uint32_t* flag_mapped, uint32_t seq, void* stream) {
  constexpr uint32_t kRec = sizeof(DecideOut) / 4;
  if (!params_bf16 || !norm || !counters || !dst_mapped || !flag_mapped ||
      decide_word < sizeof(scaph_level_counters) / 4 || decide_word + kRec > words ||
      clear_words > decide_word)
    return SCAPH_STATUS_INVALID;
  hipLaunchKernelGGL(k_decide_publish, dim3(1), dim3(SCAPH_WAVE), 0, as_stream(stream),
                     params_bf16, norm, counters, decide_word, m_u, level, words, clear_words,
                     dst_mapped, flag_mapped, seq);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_mlp_decide_frontier(const uint16_t* params_bf16, const float* norm,
                              const uint32_t* bitmap, const uint32_t* queue,
                              const scaph_page* pages, uint32_t num_pages, uint32_t n,
                              const scaph_level_counters* counters, unsigned long long m_u,
                              uint32_t level, uint32_t* page_cnt, float* out, void* stream) {
  if (!params_bf16 || !norm || !pages || !counters || !page_cnt || !out || (!bitmap && !queue))
    return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  // bitmap: one thread per 64 vertices; list: a fixed grid reading the
  // length on the device
  const unsigned grid = bitmap ? grid_for(((size_t)n + 63) / 64, kBlock, 512) : 256u;
  hipLaunchKernelGGL(k_count_frontier, dim3(grid), dim3(kBlock), 0, st, bitmap, queue, pages,
                     num_pages, n, counters, page_cnt);
  SCAPH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_decide_clear, dim3(1), dim3(kBlock), 0, st, params_bf16, norm, page_cnt,
                     pages, num_pages, counters, m_u, level, reinterpret_cast<DecideOut*>(out));
  SCAPH_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
