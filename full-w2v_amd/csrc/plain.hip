// FP32 "plain" SGNS update kernel for gfx950: the reference method
// (FULL-W2V-kernels.cu:106-250) re-designed around one wave64 per sentence.
//
//  * one wave owns one sentence; a block holds independent waves (no __syncthreads);
//  * the context window lives in an LDS ring of 2W+1 rows per wave: each context row
//    is read from HBM once when it enters the window and its accumulated delta is
//    written back once when it leaves (lifetime reuse, kernels.cu:117-120,139-145,217-249);
//  * rows are held as (load-time value U0, accumulated delta D) so the flushed delta
//    is the exact FP32 sum of its per-pair terms, not a difference of rounded rows;
//  * every target row (positive and each negative) is trained against the whole
//    window in registers and its delta is added back with one atomic per element
//    (independence of negative samples, kernels.cu:151-215);
//  * the per-pair dot product is a wave reduction on DPP + permlane swaps;
//  * sentences are pulled from a device work queue by a persistent grid, so uneven
//    sentence lengths do not leave the tail of the launch idle.
//
// Row layout: element j of row r at r*P + j with P a multiple of 64; lane l holds
// elements l + 64*e (e < E = P/64), so every row access is a 256-byte coalesced chunk.
#include "common.h"

namespace w2v {

__device__ __forceinline__ float sgns_grad(float f, bool positive, float alpha, const UpdateArgs& a) {
  if (a.sig_exact) {
    return positive ? alpha / (1.0f + expf(f)) : -alpha / (1.0f + expf(-f));
  }
  const int idx = uniform_i(sigmoid_index(f));
  return positive ? alpha * a.nexp_t[idx] : -alpha * a.exp_t[idx];
}

__device__ __forceinline__ void det_fence(const UpdateArgs& a) {
  if (a.fence) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
}

template <int E>
__global__ void __launch_bounds__(256) sgns_plain_kernel(UpdateArgs a) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  const int W = a.W, K = a.K, P = a.P, d = a.d;
  const int R = 2 * W + 1;
  float* U0s = lds + (size_t)wib * 2 * R * P;
  float* Ds = U0s + (size_t)R * P;
  const int cpb = K > 0 ? 64 / K : 0;  // centres whose negatives fit in one VGPR
  unsigned long long words = 0, sents = 0;

  for (;;) {
    unsigned s32 = 0;
    if (lane == 0) s32 = atomicAdd(a.queue, 1u);
    const int64_t s = (int64_t)(unsigned)uniform_i((int)s32);
    if (s >= a.n_sent) break;

    int64_t base;
    int L;
    if (a.offsets) {
      base = a.offsets[s];
      L = (int)(a.offsets[s + 1] - base);
    } else {
      base = s * (int64_t)a.msl;
      L = a.length[s];
    }
    if (L <= 0) continue;
    float alpha;
    int64_t raw = 0;
    if (a.alpha) {
      alpha = a.alpha[s];
    } else {
      // Train_SGNS learning rate (driver.cu:188-189) at this sentence's corpus position
      raw = a.raw_base + a.raw_offsets[s];
      const uint64_t wca = a.epoch_words + (uint64_t)raw;
      alpha = a.alpha0 * (1.0f - (float)wca / (float)a.total_words);
      if (alpha < a.alpha0 * 0.0001f) alpha = a.alpha0 * 0.0001f;
    }
    words += (unsigned long long)L;
    sents += 1;

    // entering-token stream, 64 tokens per VGPR chunk
    int chunk0 = 0;
    int tokbuf = (lane < L) ? a.tokens[base + lane] : 0;
    int ring_tok = 0;  // lane r holds the token id of ring slot r
    int negbuf = 0, nw0 = -(1 << 30);

    auto load_row = [&](int slot, int tok) {
      const float* src = a.syn0 + (size_t)tok * P;
      float* u0 = U0s + (size_t)slot * P;
      float* dd = Ds + (size_t)slot * P;
#pragma unroll
      for (int e = 0; e < E; e++) {
        float v = src[e * 64 + lane];
        if (a.rd0) v += a.rd0[(size_t)tok * P + e * 64 + lane];
        u0[e * 64 + lane] = v;
        dd[e * 64 + lane] = 0.0f;
      }
    };
    auto flush_row = [&](int slot, int tok) {
      float* dst = a.out0 + (size_t)tok * P;
      const float* dd = Ds + (size_t)slot * P;
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int j = e * 64 + lane;
        if (j < d) atomicAdd(dst + j, dd[j]);
      }
    };
    auto negative = [&](int w, int k) -> int {
      if (w >= nw0 + cpb) {
        nw0 = w;
        const int cw = nw0 + lane / K, kk = lane % K;
        int v = 0;
        if (lane < cpb * K && cw < L) {
          if (a.neg) {
            v = a.neg[(base + cw) * K + kk];
          } else {
            const uint64_t x = key2(a.mixed_seed, pos_key(a.epoch, raw + cw), (uint64_t)kk + 1);
            const uint64_t idx = ((x >> 32) * (uint64_t)a.n_unigram) >> 32;
            v = a.unigram[idx];
          }
        }
        negbuf = v;
      }
      return readlane_i(negbuf, (w - nw0) * K + k);
    };

    // Hogwild mode loads the next target row and the next entering context row one
    // step ahead; deterministic mode loads in order after its fences.
    const bool pipe = !a.fence;
    auto fetch = [&](const float* m, const float* dm, int row, float* dst) {
#pragma unroll
      for (int e = 0; e < E; e++) {
        float v = m[(size_t)row * P + e * 64 + lane];
        if (dm) v += dm[(size_t)row * P + e * 64 + lane];
        dst[e] = v;
      }
    };
    auto entering_token = [&](int pos) -> int {
      if (pos - chunk0 >= 64) {
        chunk0 += 64;
        tokbuf = (chunk0 + lane < L) ? a.tokens[base + chunk0 + lane] : 0;
      }
      return readlane_i(tokbuf, pos - chunk0);
    };

    // preload positions 0..W (kernels.cu:139-145)
    const int npre = L < W + 1 ? L : W + 1;
    for (int p = 0; p < npre; p++) {
      const int tok = readlane_i(tokbuf, p);
      ring_tok = (lane == p) ? tok : ring_tok;
      load_row(p, tok);
    }
    float nT[E], nE[E];  // prefetched next target row / next entering row
    int t_next = readlane_i(ring_tok, 0);
    if (pipe) fetch(a.syn1, a.rd1, t_next, nT);

    int cs = 0;  // ring slot of the centre
    for (int w = 0; w < L; w++) {
      const int lo = w - W < 0 ? 0 : w - W;
      const int hi = w + W > L - 1 ? L - 1 : w + W;
      const int lo_slot0 = cs - (w - lo) < 0 ? cs - (w - lo) + R : cs - (w - lo);
      const int enter_pos = w + W + 1;
      int enter_tok = 0;
      if (enter_pos < L) {
        enter_tok = entering_token(enter_pos);
        if (pipe) fetch(a.syn0, a.rd0, enter_tok, nE);
      }
      for (int k = 0; k <= K; k++) {
        const int t = t_next;
        float T0[E], DT[E];
        if (pipe) {
#pragma unroll
          for (int e = 0; e < E; e++) T0[e] = nT[e];
        } else {
          fetch(a.syn1, a.rd1, t, T0);
        }
#pragma unroll
        for (int e = 0; e < E; e++) DT[e] = 0.0f;
        // next target: the next negative of this centre, or the next centre's word
        if (k < K) t_next = negative(w, k);
        else if (w + 1 < L) t_next = readlane_i(ring_tok, cs + 1 == R ? 0 : cs + 1);
        if (pipe && (k < K || w + 1 < L)) fetch(a.syn1, a.rd1, t_next, nT);
        int slot = lo_slot0;
        for (int c = lo; c <= hi; c++, slot = (slot + 1 == R) ? 0 : slot + 1) {
          if (c == w) continue;
          float* u0 = U0s + (size_t)slot * P;
          float* dd = Ds + (size_t)slot * P;
          float uu[E], tt[E], dv[E];
          float part = 0.0f;
#pragma unroll
          for (int e = 0; e < E; e++) {
            dv[e] = dd[e * 64 + lane];
            uu[e] = u0[e * 64 + lane] + dv[e];
            tt[e] = T0[e] + DT[e];
            part = fmaf(uu[e], tt[e], part);
          }
          const float f = wave_sum(part);
          const float g = sgns_grad(f, k == 0, alpha, a);
#pragma unroll
          for (int e = 0; e < E; e++) {
            dd[e * 64 + lane] = fmaf(g, tt[e], dv[e]);  // context and target move together
            DT[e] = fmaf(g, uu[e], DT[e]);
          }
        }
        float* tdst = a.out1 + (size_t)t * P;
#pragma unroll
        for (int e = 0; e < E; e++) {
          const int j = e * 64 + lane;
          if (j < d) atomicAdd(tdst + j, DT[e]);
        }
        det_fence(a);
      }
      // ring advance: flush the leaving context, then load the entering one into the
      // same slot (kernels.cu:217-238)
      const int leave = w - W, enter = w + W + 1;
      int slot = cs + W + 1;
      if (slot >= R) slot -= R;
      if (leave >= 0) {
        flush_row(slot, readlane_i(ring_tok, slot));
        det_fence(a);
      }
      if (enter < L) {
        ring_tok = (lane == slot) ? enter_tok : ring_tok;
        if (pipe) {
          float* u0 = U0s + (size_t)slot * P;
          float* dd = Ds + (size_t)slot * P;
#pragma unroll
          for (int e = 0; e < E; e++) {
            u0[e * 64 + lane] = nE[e];
            dd[e * 64 + lane] = 0.0f;
          }
        } else {
          load_row(slot, enter_tok);
        }
      }
      cs = (cs + 1 == R) ? 0 : cs + 1;
    }
    // final flush of the contexts still in the window (kernels.cu:245-249)
    for (int p = (L - W > 0 ? L - W : 0); p < L; p++) {
      const int slot = p % R;
      flush_row(slot, readlane_i(ring_tok, slot));
    }
    det_fence(a);
  }
  if (lane == 0 && a.counters) {
    atomicAdd(a.counters + 0, words);
    atomicAdd(a.counters + 1, sents);
  }
}

// LDS bytes per wave of the plain kernel.
size_t plain_lds_per_wave(int W, int P) { return (size_t)2 * (2 * W + 1) * P * sizeof(float); }

template <int E>
static hipError_t launch_plain_e(const UpdateArgs& a, int waves_per_block, int blocks,
                                 hipStream_t st) {
  const size_t lds = plain_lds_per_wave(a.W, a.P) * waves_per_block;
  hipError_t err = hipFuncSetAttribute((const void*)sgns_plain_kernel<E>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(sgns_plain_kernel<E>, dim3(blocks), dim3(64 * waves_per_block), lds, st, a);
  return hipGetLastError();
}

template <int E>
static hipError_t occupancy_plain_e(int waves_per_block, size_t lds, int* blocks_per_cu) {
  hipError_t err = hipFuncSetAttribute((const void*)sgns_plain_kernel<E>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (err != hipSuccess) return err;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, sgns_plain_kernel<E>,
                                                      64 * waves_per_block, lds);
}

#define W2V_E_CASES(M) M(1) M(2) M(3) M(4) M(5) M(6) M(7) M(8) M(9) M(10) M(11) M(12) M(13) M(14) M(15) M(16)

hipError_t launch_plain(const UpdateArgs& a, int waves_per_block, int blocks, hipStream_t st) {
  switch (a.P / 64) {
#define CASE(E) \
  case E: return launch_plain_e<E>(a, waves_per_block, blocks, st);
    W2V_E_CASES(CASE)
#undef CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t occupancy_plain(int P, int waves_per_block, size_t lds, int* blocks_per_cu) {
  switch (P / 64) {
#define CASE(E) \
  case E: return occupancy_plain_e<E>(waves_per_block, lds, blocks_per_cu);
    W2V_E_CASES(CASE)
#undef CASE
    default: return hipErrorInvalidValue;
  }
}

}  // namespace w2v
