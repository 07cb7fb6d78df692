// Synthetic corpora and model initialisation on the device.
//
//  * cluster-Markov Zipf corpus: every word id 1..V-1 belongs to one of C clusters;
//    tokens follow a cluster Markov chain (stay probability), each drawn from the
//    cluster's Zipf-restricted distribution or the global Zipf; sentence lengths are
//    rounded log-normal, clamped to [1, max]. All draws are integer comparisons of
//    keyed 53-bit uniforms against thresholds the host computes in FP64, so the CPU
//    restatement produces the identical corpus.
//  * word counts (the vocabulary counts the unigram and subsampling tables need);
//  * syn0 initialisation of InitNet (improved_driver_functions.cpp:27-31): the LCG
//    stream from seed 1 over all V*d elements, computed in parallel by jump-ahead.
#include <hipcub/hipcub.hpp>
#include <math.h>

#include <vector>

#include "common.h"
#include "synth_tables.h"

namespace w2v {

constexpr uint64_t kSaltCluster = 0x636C7573746572ull;
constexpr uint64_t kSaltSent = 0x53454E54ull;
constexpr uint64_t kSaltLen = 0x4C454Eull;
constexpr double kT53 = 9007199254740992.0;  // 2^53



static uint64_t host_key(uint64_t seed, uint64_t a, uint64_t b) { return key2(mix64(seed), a, b); }

static uint64_t thr_of(double cum, double total) {
  double q = floor(cum / total * kT53);
  if (q >= kT53) return (uint64_t)kT53;
  return (uint64_t)q;
}

int32_t synth_cluster_of(uint64_t seed, int32_t clusters, int32_t word) {
  return (int32_t)(host_key(seed, kSaltCluster, (uint64_t)word) % (uint64_t)clusters);
}

bool synth_make_tables(int32_t V, int32_t C, double s, double stay, double in_cluster,
                       int32_t len_median, double len_sigma, int32_t len_max, uint64_t seed,
                       SynthTables& t) {
  if (V < 2 || C < 1 || len_max < 1) return false;
  t.gthr.assign(V, 0);
  double total = 0.0;
  for (int32_t r = 1; r < V; r++) total += pow((double)r, -s);
  double cum = 0.0;
  for (int32_t r = 1; r < V; r++) {
    cum += pow((double)r, -s);
    t.gthr[r] = (r == V - 1) ? (uint64_t)kT53 : thr_of(cum, total);
  }
  std::vector<int64_t> cnt(C + 1, 0), fill(C, 0);
  std::vector<int32_t> cl(V, 0);
  for (int32_t r = 1; r < V; r++) {
    cl[r] = synth_cluster_of(seed, C, r);
    cnt[cl[r] + 1]++;
  }
  t.cstart.assign(C + 1, 0);
  for (int32_t k = 0; k < C; k++) t.cstart[k + 1] = t.cstart[k] + cnt[k + 1];
  t.members.assign(V - 1, 0);
  for (int32_t r = 1; r < V; r++) t.members[t.cstart[cl[r]] + fill[cl[r]]++] = r;
  t.cthr.assign(V - 1, 0);
  for (int32_t k = 0; k < C; k++) {
    double tot = 0.0;
    for (int64_t i = t.cstart[k]; i < t.cstart[k + 1]; i++) tot += pow((double)t.members[i], -s);
    double cu = 0.0;
    for (int64_t i = t.cstart[k]; i < t.cstart[k + 1]; i++) {
      cu += pow((double)t.members[i], -s);
      t.cthr[i] = (i == t.cstart[k + 1] - 1) ? (uint64_t)kT53 : thr_of(cu, tot);
    }
  }
  t.lthr.assign(len_max + 1, 0);
  if (len_median <= 0) {
    t.lthr[len_max] = (uint64_t)kT53;
  } else {
    const double mu = log((double)len_median);
    for (int32_t L = 1; L <= len_max; L++) {
      const double z = (log(L + 0.5) - mu) / len_sigma;
      const double cdf = 0.5 * erfc(-z / sqrt(2.0));
      t.lthr[L] = (L == len_max) ? (uint64_t)kT53 : thr_of(cdf, 1.0);
    }
  }
  t.stay_thr = (uint64_t)floor(stay * kT53);
  t.in_thr = (uint64_t)floor(in_cluster * kT53);
  return true;
}

__device__ __forceinline__ int64_t lower_bound_u64(const uint64_t* __restrict__ thr, int64_t lo,
                                                   int64_t hi, uint64_t u) {
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (u < thr[mid]) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

__global__ void synth_len_kernel(int64_t n, int64_t first, uint64_t mixed_seed,
                                 const uint64_t* __restrict__ lthr, int32_t len_max,
                                 int64_t* __restrict__ len) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t u = key2(mixed_seed, kSaltLen, (uint64_t)(first + i)) >> 11;
  len[i] = lower_bound_u64(lthr, 1, (int64_t)len_max + 1, u);
}

__global__ void synth_tokens_kernel(int64_t n, int64_t first, uint64_t mixed_seed, int32_t V,
                                    int32_t C, const uint64_t* __restrict__ gthr,
                                    const int32_t* __restrict__ members,
                                    const int64_t* __restrict__ cstart,
                                    const uint64_t* __restrict__ cthr, uint64_t stay_thr,
                                    uint64_t in_thr, const int64_t* __restrict__ off,
                                    int32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t ss = key2(mixed_seed, kSaltSent, (uint64_t)(first + i));
  const uint64_t ms = mix64(ss);
  const int64_t b = off[i];
  const int32_t len = (int32_t)(off[i + 1] - b);
  int32_t cl = (int32_t)(key2(ms, 0, 2) % (uint64_t)C);
  for (int32_t t = 0; t < len; t++) {
    if (t > 0 && (key2(ms, (uint64_t)t, 1) >> 11) >= stay_thr)
      cl = (int32_t)(key2(ms, (uint64_t)t, 2) % (uint64_t)C);
    const uint64_t u2 = key2(ms, (uint64_t)t, 3) >> 11;
    const uint64_t u3 = key2(ms, (uint64_t)t, 4) >> 11;
    const int64_t c0 = cstart[cl], c1 = cstart[cl + 1];
    int32_t w;
    if (u2 < in_thr && c1 > c0) w = members[lower_bound_u64(cthr, c0, c1, u3)];
    else w = (int32_t)lower_bound_u64(gthr, 1, V, u3);
    out[b + t] = w;
  }
}

// Word counts: ids below kHead are counted in an LDS histogram per block, the tail
// with global atomics (the Zipf head would serialise on global atomics).
constexpr int kHead = 16384;

__global__ void __launch_bounds__(1024) count_kernel(const int32_t* __restrict__ tok, int64_t n,
                                                     unsigned long long* __restrict__ cn) {
  __shared__ unsigned int h[kHead];
  for (int i = threadIdx.x; i < kHead; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int32_t w = tok[i];
    if (w < kHead) atomicAdd(&h[w], 1u);
    else atomicAdd(&cn[w], 1ull);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kHead; i += blockDim.x)
    if (h[i]) atomicAdd(&cn[i], (unsigned long long)h[i]);
}

// InitNet syn0: element e (row-major over V x d) takes LCG state x_{e+1} from x_0 = 1.
// Each thread jumps to its first element with the precomputed affine powers of the
// LCG map and then steps sequentially over `chunk` elements.
__global__ void init_syn0_kernel(float* __restrict__ syn0, uint16_t* __restrict__ bsyn0, int64_t V,
                                 int32_t d, int32_t P, int32_t chunk,
                                 const uint64_t* __restrict__ jump_a,
                                 const uint64_t* __restrict__ jump_c) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = V * (int64_t)d;
  const int64_t e0 = t * chunk;
  if (e0 >= total) return;
  uint64_t x = 1;
  uint64_t steps = (uint64_t)e0;  // state before element e0 is x_{e0}
  for (int k = 0; steps; k++, steps >>= 1)
    if (steps & 1) x = jump_a[k] * x + jump_c[k];
  const int64_t e1 = e0 + chunk < total ? e0 + chunk : total;
  for (int64_t e = e0; e < e1; e++) {
    x = x * 25214903917ull + 11ull;
    const float v = (((float)(x & 0xFFFF) / 65536.f) - 0.5f) / (float)d;
    const int64_t at = (e / d) * P + (e % d);
    if (bsyn0) bsyn0[at] = __builtin_bit_cast(uint16_t, (__bf16)v);  // round to nearest even
    else syn0[at] = v;
  }
}

hipError_t launch_synth_lengths(int64_t n, int64_t first, uint64_t mixed_seed, const uint64_t* lthr,
                                int32_t len_max, int64_t* len, hipStream_t st) {
  hipLaunchKernelGGL(synth_len_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, first,
                     mixed_seed, lthr, len_max, len);
  return hipGetLastError();
}

hipError_t scan_lengths(const int64_t* len, int64_t* off, int64_t n, void* temp, size_t* temp_bytes,
                        hipStream_t st) {
  // off has n+1 entries; len[n] must be 0
  return hipcub::DeviceScan::ExclusiveSum(temp, *temp_bytes, len, off, (int)(n + 1), st);
}

hipError_t launch_synth_tokens(int64_t n, int64_t first, uint64_t mixed_seed, int32_t V, int32_t C,
                               const uint64_t* gthr, const int32_t* members, const int64_t* cstart,
                               const uint64_t* cthr, uint64_t stay_thr, uint64_t in_thr,
                               const int64_t* off, int32_t* out, hipStream_t st) {
  hipLaunchKernelGGL(synth_tokens_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, n,
                     first, mixed_seed, V, C, gthr, members, cstart, cthr, stay_thr, in_thr, off, out);
  return hipGetLastError();
}

hipError_t launch_count(const int32_t* tok, int64_t n, unsigned long long* cn, int num_cu,
                        hipStream_t st) {
  hipLaunchKernelGGL(count_kernel, dim3(num_cu * 2), dim3(1024), 0, st, tok, n, cn);
  return hipGetLastError();
}

void lcg_jump_tables(uint64_t* a, uint64_t* c) {
  // powers 2^k of the affine map x -> A x + C (mod 2^64)
  uint64_t ak = 25214903917ull, ck = 11ull;
  for (int k = 0; k < 64; k++) {
    a[k] = ak;
    c[k] = ck;
    ck = ak * ck + ck;
    ak = ak * ak;
  }
}

hipError_t launch_init_syn0(float* syn0, uint16_t* bsyn0, int64_t V, int32_t d, int32_t P,
                            const uint64_t* ja, const uint64_t* jc, hipStream_t st) {
  const int chunk = 64;
  const int64_t threads = (V * (int64_t)d + chunk - 1) / chunk;
  hipLaunchKernelGGL(init_syn0_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, syn0,
                     bsyn0, V, d, P, chunk, ja, jc);
  return hipGetLastError();
}

}  // namespace w2v
