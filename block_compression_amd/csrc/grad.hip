// Decode adjoints (BC7, BC1) and gradient endpoint refinement (BC7): one lane
// per block.  The maths is in decode_adjoint.hpp; these are the kernels and
// the C ABI.
#include <mutex>

#include "bcn_internal.hpp"
#include "decode_adjoint.hpp"

namespace bcn {

constexpr int kGradGroup = 64;

template <int FMT>  // BCN_BC7 or BCN_BC1
__global__ __launch_bounds__(kGradGroup) void endpoints_kernel(const uint32_t* __restrict__ blocks,
                                                                uint32_t n, float* __restrict__ x) {
  uint32_t b = blockIdx.x * kGradGroup + threadIdx.x;
  if (b >= n) return;
  if (FMT == BCN_BC7) {
    uint32_t in[4] = {blocks[4 * b], blocks[4 * b + 1], blocks[4 * b + 2], blocks[4 * b + 3]};
    Bc7Form f;
    bc7_parse(in, f);
#pragma unroll
    for (int e = 0; e < 6; e++)
#pragma unroll
      for (int c = 0; c < 4; c++) x[(size_t)b * 24 + e * 4 + c] = (float)f.q[e][c];
  } else {
    uint32_t in[2] = {blocks[2 * b], blocks[2 * b + 1]};
    Bc1Form f;
    bc1_parse(in, f);
#pragma unroll
    for (int e = 0; e < 2; e++)
#pragma unroll
      for (int c = 0; c < 3; c++) x[(size_t)b * 6 + e * 3 + c] = (float)f.q[e][c];
  }
}

template <int FMT>
__global__ __launch_bounds__(kGradGroup) void forward_kernel(const uint32_t* __restrict__ blocks,
                                                              const float* __restrict__ x, uint32_t n,
                                                              int smooth, float* __restrict__ texels) {
  uint32_t b = blockIdx.x * kGradGroup + threadIdx.x;
  if (b >= n) return;
  float t[16][4];
  if (FMT == BCN_BC7) {
    uint32_t in[4] = {blocks[4 * b], blocks[4 * b + 1], blocks[4 * b + 2], blocks[4 * b + 3]};
    Bc7Form f;
    bc7_parse(in, f);
    float xv[6][4];
#pragma unroll
    for (int e = 0; e < 6; e++)
#pragma unroll
      for (int c = 0; c < 4; c++) xv[e][c] = x[(size_t)b * 24 + e * 4 + c];
    bc7_forward(f, xv, smooth != 0, t);
  } else {
    uint32_t in[2] = {blocks[2 * b], blocks[2 * b + 1]};
    Bc1Form f;
    bc1_parse(in, f);
    float xv[2][3];
#pragma unroll
    for (int e = 0; e < 2; e++)
#pragma unroll
      for (int c = 0; c < 3; c++) xv[e][c] = x[(size_t)b * 6 + e * 3 + c];
    bc1_forward(f, xv, smooth != 0, t);
  }
  float4* o = reinterpret_cast<float4*>(texels + (size_t)b * 64);
#pragma unroll
  for (int k = 0; k < 16; k++) o[k] = make_float4(t[k][0], t[k][1], t[k][2], t[k][3]);
}

template <int FMT>
__global__ __launch_bounds__(kGradGroup) void backward_kernel(const uint32_t* __restrict__ blocks,
                                                               const float* __restrict__ gt, uint32_t n,
                                                               float* __restrict__ gx) {
  uint32_t b = blockIdx.x * kGradGroup + threadIdx.x;
  if (b >= n) return;
  float g[16][4];
  const float4* gi = reinterpret_cast<const float4*>(gt + (size_t)b * 64);
#pragma unroll
  for (int k = 0; k < 16; k++) {
    float4 v = gi[k];
    g[k][0] = v.x;
    g[k][1] = v.y;
    g[k][2] = v.z;
    g[k][3] = v.w;
  }
  if (FMT == BCN_BC7) {
    uint32_t in[4] = {blocks[4 * b], blocks[4 * b + 1], blocks[4 * b + 2], blocks[4 * b + 3]};
    Bc7Form f;
    bc7_parse(in, f);
    float out[6][4];
    bc7_backward(f, g, out);
#pragma unroll
    for (int e = 0; e < 6; e++)
#pragma unroll
      for (int c = 0; c < 4; c++) gx[(size_t)b * 24 + e * 4 + c] = out[e][c];
  } else {
    uint32_t in[2] = {blocks[2 * b], blocks[2 * b + 1]};
    Bc1Form f;
    bc1_parse(in, f);
    float out[2][3];
    bc1_backward(f, g, out);
#pragma unroll
    for (int e = 0; e < 2; e++)
#pragma unroll
      for (int c = 0; c < 3; c++) gx[(size_t)b * 6 + e * 3 + c] = out[e][c];
  }
}

// Texel-side adjoint (bc7_texel_backward / bc1_texel_backward): gradients
// (n, 16, 4) in, (n, 16, 4) out, one lane per block.
template <int FMT>
__global__ __launch_bounds__(kGradGroup) void texel_backward_kernel(const uint32_t* __restrict__ blocks,
                                                                     const float* __restrict__ gd, uint32_t n,
                                                                     float* __restrict__ gs) {
  uint32_t b = blockIdx.x * kGradGroup + threadIdx.x;
  if (b >= n) return;
  float g[16][4], out[16][4];
  const float4* gi = reinterpret_cast<const float4*>(gd + (size_t)b * 64);
#pragma unroll
  for (int k = 0; k < 16; k++) {
    float4 v = gi[k];
    g[k][0] = v.x;
    g[k][1] = v.y;
    g[k][2] = v.z;
    g[k][3] = v.w;
  }
  if (FMT == BCN_BC7) {
    uint32_t in[4] = {blocks[4 * b], blocks[4 * b + 1], blocks[4 * b + 2], blocks[4 * b + 3]};
    Bc7Form f;
    bc7_parse(in, f);
    bc7_texel_backward(f, g, out);
  } else {
    uint32_t in[2] = {blocks[2 * b], blocks[2 * b + 1]};
    Bc1Form f;
    bc1_parse(in, f);
    bc1_texel_backward(f, g, out);
  }
  float4* o = reinterpret_cast<float4*>(gs + (size_t)b * 64);
#pragma unroll
  for (int k = 0; k < 16; k++) o[k] = make_float4(out[k][0], out[k][1], out[k][2], out[k][3]);
}

// ---- refinement: blocks sorted by mode, four lanes per block -------------------
//
// The lanes of a quad refine one block together, lane c owning internal
// channel c: its endpoint values are the lane's parameters (with their Adam
// state), and its texels' residuals give that channel's share of the squared
// error and gradient.  Only two things cross lanes: the squared error (the
// block's score) and, when indices are reassigned, each candidate's error;
// both are summed over the quad as (c0 + c1) + (c2 + c3) with DPP, so every
// lane holds the same value and takes the same decisions.  This is the order
// bc7_refine_form sums in, so the results are bit-identical to it.
//
// Code is specialised per mode, so the blocks are first counting-sorted by
// mode (order within a mode is immaterial: a block's result depends on that
// block alone) and every wave takes 16 blocks of one mode.
//
// Workspace (u32 words): [0, 8) blocks per mode, [8, 16) scatter cursors,
// [16, 24) first sorted slot of each mode, [24, 33) first wave of each mode
// (and the total), then the permutation, one word per block.
constexpr uint32_t kRefineWsHeader = 40;
constexpr uint32_t kQuadBlocks = kGradGroup / 4;  // blocks per wave

__device__ __forceinline__ uint32_t block_mode(const uint32_t* in, uint32_t b) {
  uint32_t first = in[4 * (size_t)b] & 0xFFu;
  return first == 0 ? 8u : (uint32_t)__builtin_ctz(first);
}

// Sum over the four lanes of a quad, (l0 + l1) + (l2 + l3) in every lane.
// The lane moves are written as update_dpp with all rows and banks enabled
// and bound_ctrl set, which lets the compiler fold each into its add
// (v_add_f32 with a DPP source: one instruction per step instead of two).
__device__ __forceinline__ float quad_sum(float v) {
  float a = v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, true));
  return a + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(a), 0x4E, 0xF, 0xF, true));
}

__global__ void refine_clear_kernel(uint32_t* ws) {
  if (threadIdx.x < kRefineWsHeader) ws[threadIdx.x] = 0;
}

// Blocks per mode, one atomic per mode per wave (lanes past n vote for no mode).
__global__ __launch_bounds__(256) void refine_count_kernel(const uint32_t* __restrict__ in, uint32_t n,
                                                           uint32_t* __restrict__ ws) {
  uint32_t b = blockIdx.x * 256 + threadIdx.x;
  uint32_t m = b < n ? block_mode(in, b) : 9u;
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (uint32_t mm = 0; mm < 8; mm++) {
    uint64_t bal = __ballot(m == mm);
    if (bal != 0 && lane == (uint32_t)__ffsll((unsigned long long)bal) - 1u)
      atomicAdd(&ws[mm], (uint32_t)__popcll(bal));
  }
}

__global__ void refine_offsets_kernel(uint32_t* ws) {
  if (threadIdx.x != 0) return;
  uint32_t slot = 0, wave = 0;
  for (int m = 0; m < 8; m++) {
    ws[16 + m] = slot;
    ws[8 + m] = slot;
    ws[24 + m] = wave;
    slot += ws[m];
    wave += (ws[m] + kQuadBlocks - 1) / kQuadBlocks;
  }
  ws[32] = wave;
}

// Sorts valid blocks into the permutation (one atomic per mode per wave);
// blocks with no valid mode are finished here: they decode to zero, are
// copied unchanged, and their error is the target's energy.
__global__ __launch_bounds__(kGradGroup) void refine_scatter_kernel(
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t bw, uint32_t n,
    const uint8_t* __restrict__ target, uint32_t pitch, RefineParams rp, float* __restrict__ sse_before,
    float* __restrict__ sse_after, uint32_t* __restrict__ ws) {
  const uint32_t b = blockIdx.x * kGradGroup + threadIdx.x, lane = threadIdx.x;
  const uint32_t m = b < n ? block_mode(in, b) : 9u;
#pragma unroll
  for (uint32_t mm = 0; mm < 8; mm++) {
    uint64_t bal = __ballot(m == mm);
    if (bal == 0) continue;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)bal) - 1u;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&ws[8 + mm], (uint32_t)__popcll(bal));
    base = (uint32_t)__shfl((int)base, (int)leader);
    if (m == mm) {
      uint32_t rank = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
      ws[kRefineWsHeader + base + rank] = b;
    }
  }
  if (m != 8) return;
  uint32_t yy = b / bw, xx = b - yy * bw;
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const float* px = reinterpret_cast<const float*>(target + (size_t)(4 * yy + k / 4) * pitch) + 4 * (4 * xx + k % 4);
#pragma unroll
    for (int o = 0; o < 4; o++) {
      float t = px[o] * 255.0f;
      s += o < (int)rp.channels ? t * t : 0.0f;
    }
  }
  *reinterpret_cast<uint4*>(out + 4 * (size_t)b) =
      make_uint4(in[4 * (size_t)b], in[4 * (size_t)b + 1], in[4 * (size_t)b + 2], in[4 * (size_t)b + 3]);
  if (sse_before) sse_before[b] = s;
  if (sse_after) sse_after[b] = s;
}

// ---- per-lane pieces of the quad refinement (lane = internal channel c) ----

template <int M>
__device__ __forceinline__ int quad_bits(int c) {
  constexpr Bc7ModeInfo mi = bc7_mode_c(M);
  return c < 3 ? mi.cb : mi.ab;
}

// Endpoint bytes of this lane's channel for parameters x.
template <int M>
__device__ __forceinline__ void quad_bytes(const Bc7Form& f, int c, const float x[6], int eb[6]) {
  constexpr Bc7ModeInfo mi = bc7_mode_c(M);
  constexpr int ne = 2 * mi.ns;
  const float qm = (float)((1 << quad_bits<M>(c)) - 1);
#pragma unroll
  for (int e = 0; e < 6; e++) {
    if (e >= ne) continue;
    int q = (int)fminf(fmaxf(rintf(x[e]), 0.0f), qm);
    eb[e] = bc7_ep_byte(mi, c, q, f.p[e]);
  }
}

// bc7_refine_pass for this lane's channel; returns the quad-summed error.
// t: this lane's target values, texel k at t[k * kGradGroup] (an LDS column).
template <int M, bool GRAD>
__device__ __forceinline__ float quad_pass(const Bc7Form& f, uint32_t pattern, int c, bool scored,
                                           const float* t, const float x[6], float g[6]) {
  constexpr Bc7ModeInfo mi = bc7_mode_c(M);
  constexpr int ne = 2 * mi.ns;
  int eb[6] = {0, 0, 0, 0, 0, 0};
  quad_bytes<M>(f, c, x, eb);
  float acc[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  float sse = 0.0f;
  const uint32_t pat = fresh(pattern);
  const float* tl = fresh_ptr(t);
  Bc7Form fl = f;
  fl.idx[0] = fresh(f.idx[0]);
  fl.idx[1] = fresh(f.idx[1]);
  fl.idx2[0] = fresh(f.idx2[0]);
  fl.idx2[1] = fresh(f.idx2[1]);
#pragma unroll
  for (int k = 0; k < 16; k++) {
    int s = bc7_texel_subset(pat, k);
    int wc, wa;
    bc7_texel_weights(fl, mi, k, wc, wa);
    int w = c < 3 ? wc : wa;
    int lo = s == 0 ? eb[0] : (s == 1 ? eb[2] : eb[4]);
    int hi = s == 0 ? eb[1] : (s == 1 ? eb[3] : eb[5]);
    float r = scored ? (float)interp(lo, hi, w) - tl[k * kGradGroup] : 0.0f;
    sse += r * r;
    if (!GRAD) continue;
    float glo = 2.0f * r * ((float)(64 - w) / 64.0f), ghi = 2.0f * r * ((float)w / 64.0f);
    acc[0] += s == 0 ? glo : 0.0f;
    acc[1] += s == 0 ? ghi : 0.0f;
    acc[2] += s == 1 ? glo : 0.0f;
    acc[3] += s == 1 ? ghi : 0.0f;
    acc[4] += s == 2 ? glo : 0.0f;
    acc[5] += s == 2 ? ghi : 0.0f;
  }
  if (GRAD) {
    const float slope = bc7_ep_slope(mi, c);
#pragma unroll
    for (int e = 0; e < 6; e++) g[e] = e < ne ? acc[e] * slope : 0.0f;
  }
  return quad_sum(sse);
}

// bc7_reassign_indices with the per-candidate error summed over the quad.
// Texel and candidate loops are left rolled (k and i are wave-uniform), which
// keeps this rarely run step small in code and registers.
template <int M>
__device__ __forceinline__ void quad_reassign(Bc7Form& f, uint32_t pattern, int c, bool scored,
                                              const float* t, const float x[6]) {
  constexpr Bc7ModeInfo mi = bc7_mode_c(M);
  int eb[6] = {0, 0, 0, 0, 0, 0};
  quad_bytes<M>(f, c, x, eb);
  const bool in_a = mi.ib2 == 0 || (f.isb ? c == 3 : c < 3);
  const bool in_b = mi.ib2 != 0 && (f.isb ? c < 3 : c == 3);
#pragma unroll 1
  for (int k = 0; k < 16; k++) {
    int s = bc7_texel_subset(pattern, k);
    int lo = s == 0 ? eb[0] : (s == 1 ? eb[2] : eb[4]);
    int hi = s == 0 ? eb[1] : (s == 1 ? eb[3] : eb[5]);
    const float tk = t[k * kGradGroup];
    float best_a = 3.4e38f;
    uint32_t ia = 0;
#pragma unroll 4
    for (int i = 0; i < (1 << mi.ib); i++) {
      float d = (float)interp(lo, hi, bptc_weight_rt(mi.ib, i)) - tk;
      float err = quad_sum((in_a && scored) ? d * d : 0.0f);
      if (err < best_a) {
        best_a = err;
        ia = (uint32_t)i;
      }
    }
    set_nib(f.idx, k, ia);
    if (mi.ib2) {
      float best_b = 3.4e38f;
      uint32_t ib = 0;
#pragma unroll 4
      for (int i = 0; i < (1 << mi.ib2); i++) {
        float d = (float)interp(lo, hi, bptc_weight_rt(mi.ib2, i)) - tk;
        float err = quad_sum((in_b && scored) ? d * d : 0.0f);
        if (err < best_b) {
          best_b = err;
          ib = (uint32_t)i;
        }
      }
      set_nib(f.idx2, k, ib);
    }
  }
}

template <int M>
__device__ __forceinline__ void quad_flip_pbits(Bc7Form& f, uint32_t pattern, int c, bool scored,
                                                const float* t, const float x[6]) {
  constexpr Bc7ModeInfo mi = bc7_mode_c(M);
  if (mi.pbit == 0) return;
  float g[6];
  float cur = quad_pass<M, false>(f, pattern, c, scored, t, x, g);
  constexpr int slots = mi.pbit == 1 ? 2 * mi.ns : mi.ns;
#pragma unroll
  for (int i = 0; i < slots; i++) {
    const int e0 = mi.pbit == 1 ? i : 2 * i, e1 = mi.pbit == 1 ? i : 2 * i + 1;
    f.p[e0] ^= 1;
    if (e1 != e0) f.p[e1] ^= 1;
    float sse = quad_pass<M, false>(f, pattern, c, scored, t, x, g);
    if (sse < cur) {
      cur = sse;
    } else {
      f.p[e0] ^= 1;
      if (e1 != e0) f.p[e1] ^= 1;
    }
  }
}

// bc7_refine_form over a quad.  Returns the kept error; out[] is meaningful
// in every lane (all lanes assemble the same block).
template <int M>
__device__ __forceinline__ float refine_quad(const uint32_t blk[4], int c, const float* t,
                                             const RefineParams& rp, uint32_t out[4], float* sse0) {
  constexpr Bc7ModeInfo mi = bc7_mode_c(M);
  constexpr int ne = 2 * mi.ns;
  Bc7Form f;
  bc7_parse(blk, f);
  out[0] = blk[0];
  out[1] = blk[1];
  out[2] = blk[2];
  out[3] = blk[3];
  const uint32_t pattern = bc7_pattern(mi, f.part);
  const int o = f.rot == 0 ? c : (c == 3 ? (int)f.rot - 1 : (c == (int)f.rot - 1 ? 3 : c));
  const bool scored = o < (int)rp.channels;
  const bool has = quad_bits<M>(c) != 0;
  const float qm = (float)((1 << quad_bits<M>(c)) - 1);
  float x[6], m[6], v[6], g[6];
  int bq[6];
#pragma unroll
  for (int e = 0; e < 6; e++) {
    int q = c == 0 ? f.q[e][0] : (c == 1 ? f.q[e][1] : (c == 2 ? f.q[e][2] : f.q[e][3]));
    x[e] = (float)q;
    bq[e] = q;
    m[e] = 0.0f;
    v[e] = 0.0f;
  }
  float best = quad_pass<M, true>(f, pattern, c, scored, t, x, g);
  *sse0 = best;
  const float best0 = best;
  uint32_t bidx[2] = {f.idx[0], f.idx[1]}, bidx2[2] = {f.idx2[0], f.idx2[1]};
  uint32_t bp = 0;
#pragma unroll
  for (int e = 0; e < 6; e++) bp |= (uint32_t)f.p[e] << e;
  float b1t = 1.0f, b2t = 1.0f;
  for (uint32_t it = 0; it < rp.steps; it++) {
    b1t *= rp.beta1;
    b2t *= rp.beta2;
    const float c1 = 1.0f / (1.0f - b1t), c2 = 1.0f / (1.0f - b2t);
#pragma unroll
    for (int e = 0; e < ne; e++) {
      if (!has) continue;
      m[e] = rp.beta1 * m[e] + (1.0f - rp.beta1) * g[e];
      v[e] = rp.beta2 * v[e] + (1.0f - rp.beta2) * g[e] * g[e];
      float nx = x[e] - rp.lr * (m[e] * c1) / (sqrtf(v[e] * c2) + rp.eps);
      x[e] = fminf(fmaxf(nx, -0.49f), qm + 0.49f);
    }
    if (rp.reassign && (it + 1) % rp.reassign == 0) quad_reassign<M>(f, pattern, c, scored, t, x);
    if (mi.pbit && rp.pbits && (it + 1) % rp.pbits == 0) quad_flip_pbits<M>(f, pattern, c, scored, t, x);
    float sse = quad_pass<M, true>(f, pattern, c, scored, t, x, g);
    if (sse < best) {
      best = sse;
#pragma unroll
      for (int e = 0; e < ne; e++) bq[e] = (int)fminf(fmaxf(rintf(x[e]), 0.0f), qm);
      bidx[0] = f.idx[0];
      bidx[1] = f.idx[1];
      bidx2[0] = f.idx2[0];
      bidx2[1] = f.idx2[1];
      bp = 0;
#pragma unroll
      for (int e = 0; e < 6; e++) bp |= (uint32_t)f.p[e] << e;
    }
  }
  // Every lane gathers the quad's channels (all lanes of a quad share the
  // decisions above, so the whole quad is on the same path here).
  const int base = (int)(threadIdx.x & ~3u);
#pragma unroll
  for (int e = 0; e < 6; e++)
#pragma unroll
    for (int cc = 0; cc < 4; cc++) f.q[e][cc] = __shfl(bq[e], base + cc);
  if (best < best0) {
    f.idx[0] = bidx[0];
    f.idx[1] = bidx[1];
    f.idx2[0] = bidx2[0];
    f.idx2[1] = bidx2[1];
#pragma unroll
    for (int e = 0; e < 6; e++) f.p[e] = (int)((bp >> e) & 1u);
    bc7_normalize_anchors(f);
    bc7_pack(f, out);
  }
  return best;
}

// Wave w serves mode m with ws[24 + m] <= w < ws[24 + m + 1] (16 blocks per
// wave); waves past the total (the grid is sized for the worst case) exit at
// once.  A quad past its mode's last block refines that mode's first block
// again and stores nothing, so that all 64 lanes stay active for the DPP.
__global__ __launch_bounds__(kGradGroup) __attribute__((amdgpu_waves_per_eu(3, 3))) void refine_sorted_kernel(
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t bw,
    const uint8_t* __restrict__ target, uint32_t pitch, RefineParams rp, float* __restrict__ sse_before,
    float* __restrict__ sse_after, const uint32_t* __restrict__ ws) {
  __shared__ float tq[16 * kGradGroup];
  const uint32_t w = blockIdx.x;
  if (w >= ws[32]) return;
  uint32_t m = 0;
#pragma unroll
  for (int i = 1; i < 8; i++) m = w >= ws[24 + i] ? (uint32_t)i : m;
  const uint32_t slot = (w - ws[24 + m]) * kQuadBlocks + (threadIdx.x >> 2);
  const bool active = slot < ws[m];
  const uint32_t b = ws[kRefineWsHeader + ws[16 + m] + (active ? slot : 0u)];
  const int c = (int)(threadIdx.x & 3u);
  uint32_t blk[4] = {in[4 * (size_t)b], in[4 * (size_t)b + 1], in[4 * (size_t)b + 2], in[4 * (size_t)b + 3]};
  // This lane's output channel of each texel.  The rotation is read from the
  // block for modes 4/5 (bits mode+1 .. mode+2); other modes have none.
  uint32_t rot = 0;
  if (m == 4 || m == 5) rot = (blk[0] >> (m + 1)) & 3u;
  const int o = rot == 0 ? c : (c == 3 ? (int)rot - 1 : (c == (int)rot - 1 ? 3 : c));
  const uint32_t yy = b / bw, xx = b - yy * bw;
  float* t = tq + threadIdx.x;  // this lane's column: texel k at t[k * 64]
#pragma unroll
  for (int k = 0; k < 16; k++)
    t[k * kGradGroup] =
        reinterpret_cast<const float*>(target + (size_t)(4 * yy + k / 4) * pitch)[4 * (4 * xx + k % 4) + o] * 255.0f;
  uint32_t res[4];
  float s0 = 0.0f, s1 = 0.0f;
  switch (m) {
    case 0: s1 = refine_quad<0>(blk, c, t, rp, res, &s0); break;
    case 1: s1 = refine_quad<1>(blk, c, t, rp, res, &s0); break;
    case 2: s1 = refine_quad<2>(blk, c, t, rp, res, &s0); break;
    case 3: s1 = refine_quad<3>(blk, c, t, rp, res, &s0); break;
    case 4: s1 = refine_quad<4>(blk, c, t, rp, res, &s0); break;
    case 5: s1 = refine_quad<5>(blk, c, t, rp, res, &s0); break;
    case 6: s1 = refine_quad<6>(blk, c, t, rp, res, &s0); break;
    default: s1 = refine_quad<7>(blk, c, t, rp, res, &s0); break;
  }
  if (active && c == 0) {
    *reinterpret_cast<uint4*>(out + 4 * (size_t)b) = make_uint4(res[0], res[1], res[2], res[3]);
    if (sse_before) sse_before[b] = s0;
    if (sse_after) sse_after[b] = s1;
  }
}

}  // namespace bcn

namespace {

int check_launch() { return hipGetLastError() == hipSuccess ? BCN_OK : BCN_ERR_HIP; }

bool fmt_ok(int fmt) { return fmt == BCN_BC7 || fmt == BCN_BC1; }

// Argument checks shared by bcn_bc7_refine and bcn_bc7_refine_noalloc; no
// device work.
int refine_check(const void* blocks_in, const void* blocks_out, uint32_t width, uint32_t height,
                 const float* target, uint32_t target_pitch, const bcn_refine_params* params,
                 const float* sse_before, const float* sse_after) {
  if (!params || width % 4 || height % 4) return BCN_ERR_INVALID_ARGUMENT;
  if (params->channels != 3 && params->channels != 4) return BCN_ERR_INVALID_ARGUMENT;
  if (width == 0 || height == 0) return BCN_OK;
  if (!blocks_in || !blocks_out || !target) return BCN_ERR_INVALID_ARGUMENT;
  if ((uint64_t)target_pitch < (uint64_t)width * 16) return BCN_ERR_INVALID_ARGUMENT;
  if ((((uintptr_t)blocks_in | (uintptr_t)blocks_out | (uintptr_t)target | target_pitch) & 15u) ||
      ((uintptr_t)sse_before & 3u) || ((uintptr_t)sse_after & 3u))
    return BCN_ERR_MISALIGNED;
  return BCN_OK;
}

// The refinement's launches on one workspace: clear the header, count blocks
// per mode, offsets, scatter (and refine the blocks no sorted wave takes),
// then the sorted, mode-specialised refinement.  Every word the later
// kernels read is written by an earlier one in this call.
int refine_launch(const void* blocks_in, void* blocks_out, uint32_t width, uint32_t height,
                  const float* target, uint32_t target_pitch, const bcn_refine_params& params,
                  float* sse_before, float* sse_after, uint32_t* ws, hipStream_t s) {
  const uint32_t bw = width / 4, nblocks = bw * (height / 4);
  bcn::RefineParams rp;
  rp.steps = params.steps;
  rp.lr = params.lr;
  rp.beta1 = params.beta1;
  rp.beta2 = params.beta2;
  rp.eps = params.eps;
  rp.channels = params.channels;
  rp.reassign = params.reassign;
  rp.pbits = params.pbits;
  const uint32_t* in = static_cast<const uint32_t*>(blocks_in);
  uint32_t* out = static_cast<uint32_t*>(blocks_out);
  const uint8_t* tg = reinterpret_cast<const uint8_t*>(target);
  const uint32_t waves = (nblocks + bcn::kGradGroup - 1) / bcn::kGradGroup;
  hipLaunchKernelGGL(bcn::refine_clear_kernel, dim3(1), dim3(64), 0, s, ws);
  hipLaunchKernelGGL(bcn::refine_count_kernel, dim3((nblocks + 255) / 256), dim3(256), 0, s, in, nblocks, ws);
  hipLaunchKernelGGL(bcn::refine_offsets_kernel, dim3(1), dim3(64), 0, s, ws);
  hipLaunchKernelGGL(bcn::refine_scatter_kernel, dim3(waves), dim3(bcn::kGradGroup), 0, s, in, out, bw, nblocks,
                     tg, target_pitch, rp, sse_before, sse_after, ws);
  // Sixteen blocks per wave, at most one partial wave per mode.
  const uint32_t qwaves = (nblocks + bcn::kQuadBlocks - 1) / bcn::kQuadBlocks + 8;
  hipLaunchKernelGGL(bcn::refine_sorted_kernel, dim3(qwaves), dim3(bcn::kGradGroup), 0, s, in, out, bw, tg,
                     target_pitch, rp, sse_before, sse_after, ws);
  return check_launch();
}

// ---- the allocation-free refinement's scratch --------------------------------
//
// A fixed device array replaces the caller's workspace: header plus one word
// per block for up to kNoAllocMaxBlocks blocks (4096 x 4096 texels), 4 MiB
// per GPU, held until the process ends.  In host code `scratch` names a
// host-side stand-in, not device memory: only the address that
// hipGetSymbolAddress returns for the current device is passed to kernels.
constexpr uint32_t kNoAllocMaxBlocks = BCN_NOALLOC_MAX_BLOCKS;
__device__ __align__(256) uint32_t scratch[bcn::kRefineWsHeader + kNoAllocMaxBlocks];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices];

// The current device's copy of `scratch`, looked up once per device.
int noalloc_scratch(uint32_t** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return BCN_ERR_NO_DEVICE;
  if (dev >= kMaxDevices) return BCN_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return BCN_ERR_HIP;
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return BCN_OK;
}

}  // namespace

extern "C" {

int bcn_decode_endpoints(bcn_variant variant, const void* blocks, uint32_t nblocks, float* x,
                         void* hip_stream) {
  if (!fmt_ok(variant) || (nblocks && (!blocks || !x))) return BCN_ERR_INVALID_ARGUMENT;
  if (((uintptr_t)blocks & 3u) || ((uintptr_t)x & 3u)) return BCN_ERR_MISALIGNED;
  if (nblocks == 0) return BCN_OK;
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  dim3 grid((nblocks + bcn::kGradGroup - 1) / bcn::kGradGroup), block(bcn::kGradGroup);
  const uint32_t* bp = static_cast<const uint32_t*>(blocks);
  if (variant == BCN_BC7) hipLaunchKernelGGL(bcn::endpoints_kernel<BCN_BC7>, grid, block, 0, s, bp, nblocks, x);
  else hipLaunchKernelGGL(bcn::endpoints_kernel<BCN_BC1>, grid, block, 0, s, bp, nblocks, x);
  return check_launch();
}

int bcn_decode_with_endpoints(bcn_variant variant, const void* blocks, const float* x, uint32_t nblocks,
                              int smooth, float* texels, void* hip_stream) {
  if (!fmt_ok(variant) || (nblocks && (!blocks || !x || !texels))) return BCN_ERR_INVALID_ARGUMENT;
  if (((uintptr_t)blocks & 3u) || ((uintptr_t)x & 3u) || ((uintptr_t)texels & 15u)) return BCN_ERR_MISALIGNED;
  if (nblocks == 0) return BCN_OK;
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  dim3 grid((nblocks + bcn::kGradGroup - 1) / bcn::kGradGroup), block(bcn::kGradGroup);
  const uint32_t* bp = static_cast<const uint32_t*>(blocks);
  if (variant == BCN_BC7)
    hipLaunchKernelGGL(bcn::forward_kernel<BCN_BC7>, grid, block, 0, s, bp, x, nblocks, smooth, texels);
  else
    hipLaunchKernelGGL(bcn::forward_kernel<BCN_BC1>, grid, block, 0, s, bp, x, nblocks, smooth, texels);
  return check_launch();
}

int bcn_decode_with_endpoints_backward(bcn_variant variant, const void* blocks, const float* grad_texels,
                                       uint32_t nblocks, float* grad_x, void* hip_stream) {
  if (!fmt_ok(variant) || (nblocks && (!blocks || !grad_texels || !grad_x))) return BCN_ERR_INVALID_ARGUMENT;
  if (((uintptr_t)blocks & 3u) || ((uintptr_t)grad_x & 3u) || ((uintptr_t)grad_texels & 15u))
    return BCN_ERR_MISALIGNED;
  if (nblocks == 0) return BCN_OK;
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  dim3 grid((nblocks + bcn::kGradGroup - 1) / bcn::kGradGroup), block(bcn::kGradGroup);
  const uint32_t* bp = static_cast<const uint32_t*>(blocks);
  if (variant == BCN_BC7)
    hipLaunchKernelGGL(bcn::backward_kernel<BCN_BC7>, grid, block, 0, s, bp, grad_texels, nblocks, grad_x);
  else
    hipLaunchKernelGGL(bcn::backward_kernel<BCN_BC1>, grid, block, 0, s, bp, grad_texels, nblocks, grad_x);
  return check_launch();
}

int bcn_decode_texels_backward(bcn_variant variant, const void* blocks, const float* grad_texels,
                               uint32_t nblocks, float* grad_source, void* hip_stream) {
  if (!fmt_ok(variant) || (nblocks && (!blocks || !grad_texels || !grad_source)))
    return BCN_ERR_INVALID_ARGUMENT;
  if (((uintptr_t)blocks & 3u) || ((uintptr_t)grad_texels & 15u) || ((uintptr_t)grad_source & 15u))
    return BCN_ERR_MISALIGNED;
  if (nblocks == 0) return BCN_OK;
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  dim3 grid((nblocks + bcn::kGradGroup - 1) / bcn::kGradGroup), block(bcn::kGradGroup);
  const uint32_t* bp = static_cast<const uint32_t*>(blocks);
  if (variant == BCN_BC7)
    hipLaunchKernelGGL(bcn::texel_backward_kernel<BCN_BC7>, grid, block, 0, s, bp, grad_texels, nblocks,
                       grad_source);
  else
    hipLaunchKernelGGL(bcn::texel_backward_kernel<BCN_BC1>, grid, block, 0, s, bp, grad_texels, nblocks,
                       grad_source);
  return check_launch();
}

size_t bcn_bc7_refine_workspace_size(uint32_t nblocks) {
  return ((size_t)bcn::kRefineWsHeader + nblocks) * sizeof(uint32_t);
}

int bcn_bc7_refine(const void* blocks_in, void* blocks_out, uint32_t width, uint32_t height,
                   const float* target, uint32_t target_pitch, const bcn_refine_params* params,
                   float* sse_before, float* sse_after, void* workspace, size_t workspace_size,
                   void* hip_stream) {
  int rc = refine_check(blocks_in, blocks_out, width, height, target, target_pitch, params, sse_before,
                        sse_after);
  if (rc != BCN_OK || width == 0 || height == 0) return rc;
  const uint32_t nblocks = (width / 4) * (height / 4);
  if (!workspace) return BCN_ERR_INVALID_ARGUMENT;
  if (workspace_size < bcn_bc7_refine_workspace_size(nblocks)) return BCN_ERR_BUFFER_TOO_SMALL;
  if ((uintptr_t)workspace & 3u) return BCN_ERR_MISALIGNED;
  return refine_launch(blocks_in, blocks_out, width, height, target, target_pitch, *params, sse_before,
                       sse_after, static_cast<uint32_t*>(workspace), static_cast<hipStream_t>(hip_stream));
}

int bcn_bc7_refine_noalloc(const void* blocks_in, void* blocks_out, uint32_t width, uint32_t height,
                           const float* target, uint32_t target_pitch, const bcn_refine_params* params,
                           float* sse_before, float* sse_after, void* hip_stream) {
  if (params && width % 4 == 0 && height % 4 == 0 &&
      (uint64_t)(width / 4) * (height / 4) > kNoAllocMaxBlocks)
    return BCN_ERR_TOO_LARGE;
  int rc = refine_check(blocks_in, blocks_out, width, height, target, target_pitch, params, sse_before,
                        sse_after);
  if (rc != BCN_OK || width == 0 || height == 0) return rc;
  uint32_t* ws = nullptr;
  rc = noalloc_scratch(&ws);
  if (rc != BCN_OK) return rc;
  return refine_launch(blocks_in, blocks_out, width, height, target, target_pitch, *params, sse_before,
                       sse_after, ws, static_cast<hipStream_t>(hip_stream));
}

}  // extern "C"
