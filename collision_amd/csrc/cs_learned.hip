// Training and inference of the time-of-impact proposer (no counterpart in
// the reference): the network of cs_mlp.h, trained on the GPU by a fused
// forward + backward kernel per 64-sample tile (BF16 operands, FP32
// accumulation on the matrix cores), a fixed-order gradient reduction and
// an Adam kernel.  The same seed, dataset and hyper-parameters give the same
// weights bit for bit: every sum runs in a fixed order and the dataset is
// put in a canonical order (by a hash of each sample's contents) before the
// seeded shuffles.
#include <algorithm>
#include <cmath>

#include "cs_internal.h"
#include "cs_mlp.h"

namespace cs {

// inference: xs [n][24] FP32 features -> out [n][2]; grid-stride over tiles
__global__ void __launch_bounds__(MLP_THREADS)
    k_mlp_infer(const float* __restrict__ xs, int64_t n, const uint16_t* __restrict__ wb,
                const float* __restrict__ w, float* __restrict__ out) {
  __shared__ __align__(16) uint16_t x[MT * MF_INP];
  __shared__ __align__(16) uint16_t h1[MT * MF_H];
  __shared__ __align__(16) uint16_t h2[MT * MF_H];
  __shared__ float o[MT * MF_OUT];
  const int64_t ntile = (n + MT - 1) / MT;
  for (int64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
    for (int e = threadIdx.x; e < MT * MF_INP; e += MLP_THREADS) {
      const int m = e / MF_INP, f = e % MF_INP;
      const int64_t s = t * MT + m;
      x[e] = (s < n && f < MF_IN) ? f2bf(xs[s * MF_IN + f]) : (uint16_t)0;
    }
    __syncthreads();
    tile_forward(x, h1, h2, o, wb, w);
    for (int e = threadIdx.x; e < MT * MF_OUT; e += MLP_THREADS) {
      const int64_t s = t * MT + e / MF_OUT;
      if (s < n) out[s * MF_OUT + (e % MF_OUT)] = o[e];
    }
    __syncthreads();
  }
}

// Training: one 64-sample tile per work-group, forward and backward fused;
// the tile's gradient (sum over its samples of the per-sample loss
// gradient) goes to partial[blockIdx.x], its loss sum to loss_partial.
// Sample m of the tile is idx[base + m] of the dataset (idx == nullptr:
// sample base + m).
//   loss = softplus(z) - y z + y (sigmoid(s) - t)^2   (z logit, s raw seed)
// Backward in BF16 operands with FP32 accumulation: dZ3 (64 x 2, padded to
// 16 columns), dZ2 = (dZ3 W3) o [h2 > 0], dZ1 = (dZ2 W2) o [h1 > 0], and the
// weight gradients dW = dZ^T H as matrix-core products over the 64 samples.
__global__ void __launch_bounds__(MLP_THREADS)
    k_mlp_train_tile(const float* __restrict__ xs, const float* __restrict__ ex,
                     const float* __restrict__ toi, const uint32_t* __restrict__ idx, int n,
                     const uint16_t* __restrict__ wb, const float* __restrict__ w,
                     float* __restrict__ partial, float* __restrict__ loss_partial) {
  __shared__ __align__(16) uint16_t x[MT * MF_INP];
  __shared__ __align__(16) uint16_t h1[MT * MF_H];
  __shared__ __align__(16) uint16_t h2[MT * MF_H];
  __shared__ __align__(16) uint16_t dz[MT * MF_H];  // dZ2, then dZ1
  __shared__ __align__(16) uint16_t d3[MT * 16];    // dZ3 (columns 2..15 zero)
  __shared__ float o[MT * MF_OUT];
  __shared__ float lsum[MT];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, rq = (lane >> 4) * 4, r0 = 16 * wave;
  const int base = blockIdx.x * MT;
  for (int e = threadIdx.x; e < MT * MF_INP; e += MLP_THREADS) {
    const int m = e / MF_INP, f = e % MF_INP, s = base + m;
    const uint32_t src = (s < n) ? (idx ? idx[s] : (uint32_t)s) : 0u;
    x[e] = (s < n && f < MF_IN) ? f2bf(xs[(int64_t)src * MF_IN + f]) : (uint16_t)0;
  }
  __syncthreads();
  tile_forward(x, h1, h2, o, wb, w);
  if (threadIdx.x < MT) {
    const int m = threadIdx.x, s = base + m;
    float g0 = 0.f, g1 = 0.f, l = 0.f;
    if (s < n) {
      const uint32_t src = idx ? idx[s] : (uint32_t)s;
      const float y = ex[src], t = toi[src];
      const float z = o[m * 2], q = mlp_sigmoid(o[m * 2 + 1]);
      l = fmaxf(z, 0.f) - z * y + log1pf(__expf(-fabsf(z))) + y * (q - t) * (q - t);
      g0 = mlp_sigmoid(z) - y;
      g1 = y * 2.f * (q - t) * q * (1.f - q);
    }
    lsum[m] = l;
    for (int j = 0; j < 16; ++j) d3[m * 16 + j] = f2bf(j == 0 ? g0 : j == 1 ? g1 : 0.f);
  }
  __syncthreads();
  float* pg = partial + (int64_t)blockIdx.x * MF_NPARAM;
  // dW3 = dZ3^T H2 (16 padded outputs x 128, K = 64 samples), b3
  if (wave == 0) {
    floatx4 acc[8];
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) acc[nt] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < MT; kk += 16) {
      const bf16x4 a = ld4_strided(d3 + (kk + rq) * 16 + col, 16);
#pragma unroll
      for (int nt = 0; nt < 8; ++nt)
        acc[nt] = mfma16(a, ld4_strided(h2 + (kk + rq) * MF_H + 16 * nt + col, MF_H), acc[nt]);
    }
#pragma unroll
    for (int nt = 0; nt < 8; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int orow = rq + i;
        if (orow < MF_OUT) pg[MO_W3 + orow * MF_H + 16 * nt + col] = acc[nt][i];
      }
    if (lane < MF_OUT) {
      float sb = 0.f;
      for (int m = 0; m < MT; ++m) sb += bf2f(d3[m * 16 + lane]);
      pg[MO_B3 + lane] = sb;
    }
  }
  // dZ2 = (dZ3 W3) masked by ReLU (K = 16), this wavefront's 16 rows
  {
    floatx4 acc[8];
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) acc[nt] = floatx4{0.f, 0.f, 0.f, 0.f};
    wave_gemm<8>(d3 + r0 * 16, 16, wb + MB_W3T, 16, 16, acc);
#pragma unroll
    for (int nt = 0; nt < 8; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = r0 + rq + i, j = 16 * nt + col;
        dz[m * MF_H + j] = h2[m * MF_H + j] ? f2bf(acc[nt][i]) : (uint16_t)0;
      }
  }
  __syncthreads();
  // dW2 = dZ2^T H1 (128 x 128, K = 64): wavefront w owns rows 32w .. 32w+31
  for (int mt = 2 * wave; mt < 2 * wave + 2; ++mt) {
    floatx4 acc[8];
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) acc[nt] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < MT; kk += 16) {
      const bf16x4 a = ld4_strided(dz + (kk + rq) * MF_H + 16 * mt + col, MF_H);
#pragma unroll
      for (int nt = 0; nt < 8; ++nt)
        acc[nt] = mfma16(a, ld4_strided(h1 + (kk + rq) * MF_H + 16 * nt + col, MF_H), acc[nt]);
    }
#pragma unroll
    for (int nt = 0; nt < 8; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) pg[MO_W2 + (16 * mt + rq + i) * MF_H + 16 * nt + col] = acc[nt][i];
  }
  if (threadIdx.x < MF_H) {
    float sb = 0.f;
    for (int m = 0; m < MT; ++m) sb += bf2f(dz[m * MF_H + threadIdx.x]);
    pg[MO_B2 + threadIdx.x] = sb;
  }
  // dZ1 = (dZ2 W2) masked by ReLU (K = 128), this wavefront's 16 rows
  {
    floatx4 acc[8];
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) acc[nt] = floatx4{0.f, 0.f, 0.f, 0.f};
    wave_gemm<8>(dz + r0 * MF_H, MF_H, wb + MB_W2T, MF_H, MF_H, acc);
    __syncthreads();  // every wavefront has read dZ2 (here and for dW2, b2)
#pragma unroll
    for (int nt = 0; nt < 8; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = r0 + rq + i, j = 16 * nt + col;
        dz[m * MF_H + j] = h1[m * MF_H + j] ? f2bf(acc[nt][i]) : (uint16_t)0;
      }
  }
  __syncthreads();
  // dW1 = dZ1^T X (128 x 32 padded inputs, K = 64)
  for (int mt = 2 * wave; mt < 2 * wave + 2; ++mt) {
    floatx4 acc[2] = {floatx4{0.f, 0.f, 0.f, 0.f}, floatx4{0.f, 0.f, 0.f, 0.f}};
    for (int kk = 0; kk < MT; kk += 16) {
      const bf16x4 a = ld4_strided(dz + (kk + rq) * MF_H + 16 * mt + col, MF_H);
#pragma unroll
      for (int nt = 0; nt < 2; ++nt)
        acc[nt] = mfma16(a, ld4_strided(x + (kk + rq) * MF_INP + 16 * nt + col, MF_INP), acc[nt]);
    }
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int f = 16 * nt + col;
        if (f < MF_IN) pg[MO_W1 + (16 * mt + rq + i) * MF_IN + f] = acc[nt][i];
      }
  }
  if (threadIdx.x < MF_H) {
    float sb = 0.f;
    for (int m = 0; m < MT; ++m) sb += bf2f(dz[m * MF_H + threadIdx.x]);
    pg[MO_B1 + threadIdx.x] = sb;
  }
  if (threadIdx.x == 0) {
    float l = 0.f;
    for (int m = 0; m < MT; ++m) l += lsum[m];
    loss_partial[blockIdx.x] = l;
  }
}

// grad[p] = (sum over tiles, in tile order) / n; the loss likewise
__global__ void k_mlp_reduce(const float* __restrict__ partial, const float* __restrict__ lp,
                             int ntile, int n, float* __restrict__ grad, float* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < MF_NPARAM) {
    float s = 0.f;
    for (int t = 0; t < ntile; ++t) s += partial[(int64_t)t * MF_NPARAM + p];
    grad[p] = s / (float)n;
  }
  if (p == 0) {
    float s = 0.f;
    for (int t = 0; t < ntile; ++t) s += lp[t];
    *loss = s / (float)n;
  }
}

// Adam on the FP32 master weights; c1, c2 = the bias corrections 1 - beta^t
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, float lr, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MF_NPARAM) return;
  const float gg = g[p];
  const float mm = 0.9f * m[p] + 0.1f * gg;
  const float vv = 0.999f * v[p] + 0.001f * gg * gg;
  m[p] = mm;
  v[p] = vv;
  w[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + 1e-8f);
}

// BF16 copies of the weights for the matrix cores (layouts MB_*)
__global__ void k_pack_weights(const float* __restrict__ w, uint16_t* __restrict__ wb) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= MB_TOTAL) return;
  float v = 0.f;
  if (e < MB_W2) {
    const int o = e / MF_INP, f = e % MF_INP;
    v = f < MF_IN ? w[MO_W1 + o * MF_IN + f] : 0.f;
  } else if (e < MB_W3) {
    v = w[MO_W2 + (e - MB_W2)];
  } else if (e < MB_W2T) {
    const int o = (e - MB_W3) / MF_H, j = (e - MB_W3) % MF_H;
    v = o < MF_OUT ? w[MO_W3 + o * MF_H + j] : 0.f;
  } else if (e < MB_W3T) {
    const int i = (e - MB_W2T) / MF_H, j = (e - MB_W2T) % MF_H;  // W2t[i][j] = W2[j][i]
    v = w[MO_W2 + j * MF_H + i];
  } else {
    const int j = (e - MB_W3T) / 16, o = (e - MB_W3T) % 16;  // W3t[j][o] = W3[o][j]
    v = o < MF_OUT ? w[MO_W3 + o * MF_H + j] : 0.f;
  }
  wb[e] = f2bf(v);
}

__device__ __host__ inline uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// canonical-order key of each sample: a hash of its 26 values' bits
__global__ void k_ds_keys(const float* __restrict__ xs, const float* __restrict__ ex,
                          const float* __restrict__ toi, int n, uint64_t* __restrict__ keys,
                          uint32_t* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t h = 0x9e3779b97f4a7c15ull;
  for (int f = 0; f < MF_IN; ++f) h = mix64(h ^ __float_as_uint(xs[(int64_t)i * MF_IN + f]));
  h = mix64(h ^ __float_as_uint(ex[i]));
  h = mix64(h ^ __float_as_uint(toi[i]));
  keys[i] = h;
  vals[i] = (uint32_t)i;
}

// shuffle keys of one epoch: a hash of (epoch seed, position)
__global__ void k_shuffle_keys(const uint32_t* __restrict__ order, int n, uint64_t seed,
                               uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = (uint32_t)(mix64(seed ^ mix64((uint64_t)i + 0x632be59bd9b4e019ull)) >> 32);
  vals[i] = order[i];
}

namespace {

template <typename T>
hipError_t lalloc(T** p, size_t count) {
  *p = nullptr;
  return hipMalloc((void**)p, sizeof(T) * (count ? count : 1));
}
template <typename T>
void lfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

hipError_t ensure_train(LearnedState& L, int ntile) {
  if (ntile <= L.train_tiles) return hipSuccess;
  lfree(L.gpart);
  lfree(L.lpart);
  L.train_tiles = 0;
  CS_HIP_RET(lalloc(&L.gpart, (size_t)ntile * MF_NPARAM));
  CS_HIP_RET(lalloc(&L.lpart, (size_t)ntile));
  L.train_tiles = ntile;
  return hipSuccess;
}

}  // namespace

hipError_t learned_alloc(LearnedState& L) {
  if (L.w) return hipSuccess;
  CS_HIP_RET(lalloc(&L.w, MF_NPARAM));
  CS_HIP_RET(lalloc(&L.m, MF_NPARAM));
  CS_HIP_RET(lalloc(&L.v, MF_NPARAM));
  CS_HIP_RET(lalloc(&L.grad, MF_NPARAM));
  CS_HIP_RET(lalloc(&L.wb, MB_TOTAL));
  CS_HIP_RET(lalloc(&L.ds_count, 1));
  return hipMemset(L.ds_count, 0, sizeof(unsigned long long));
}

void learned_free(LearnedState& L) {
  lfree(L.w); lfree(L.m); lfree(L.v); lfree(L.grad); lfree(L.wb);
  lfree(L.ds_count); lfree(L.dsx); lfree(L.dse); lfree(L.dst);
  lfree(L.seeds); lfree(L.list); lfree(L.count);
  lfree(L.keys); lfree(L.idx); lfree(L.tkeys); lfree(L.tidx);
  lfree(L.hist); lfree(L.part); lfree(L.gpart); lfree(L.lpart);
  L = LearnedState();
}

hipError_t learned_set_weights(LearnedState& L, hipStream_t st, const float* host_w) {
  CS_HIP_RET(learned_alloc(L));
  CS_HIP_RET(hipMemcpyAsync(L.w, host_w, sizeof(float) * MF_NPARAM, hipMemcpyHostToDevice, st));
  CS_HIP_RET(hipMemsetAsync(L.m, 0, sizeof(float) * MF_NPARAM, st));
  CS_HIP_RET(hipMemsetAsync(L.v, 0, sizeof(float) * MF_NPARAM, st));
  hipLaunchKernelGGL(k_pack_weights, dim3((MB_TOTAL + 255) / 256), dim3(256), 0, st, L.w, L.wb);
  CS_HIP_RET(hipGetLastError());
  L.adam_t = 0;
  L.ready = true;
  return hipStreamSynchronize(st);
}

hipError_t learned_get_weights(LearnedState& L, hipStream_t st, float* host_w) {
  CS_HIP_RET(hipMemcpyAsync(host_w, L.w, sizeof(float) * MF_NPARAM, hipMemcpyDeviceToHost, st));
  return hipStreamSynchronize(st);
}

// He-uniform weights (limit sqrt(6 / fan_in)) from a splitmix64 stream, zero biases
void learned_init_host(uint64_t seed, float* w) {
  uint64_t s = seed;
  auto uni = [&]() {  // uniform in [-1, 1)
    s += 0x9e3779b97f4a7c15ull;
    return (double)(mix64(s) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  };
  for (int p = 0; p < MF_NPARAM; ++p) w[p] = 0.f;
  const struct {
    int off, n, fan_in;
  } layers[3] = {{MO_W1, MF_H * MF_IN, MF_IN}, {MO_W2, MF_H * MF_H, MF_H}, {MO_W3, MF_OUT * MF_H, MF_H}};
  for (const auto& l : layers) {
    const double lim = std::sqrt(6.0 / l.fan_in);
    for (int i = 0; i < l.n; ++i) w[l.off + i] = (float)(lim * uni());
  }
}

hipError_t learned_infer(LearnedState& L, hipStream_t st, const float* dx, int64_t n, float* dout) {
  if (n <= 0) return hipSuccess;
  const int64_t ntile = (n + MT - 1) / MT;
  const unsigned grid = (unsigned)std::min<int64_t>(ntile, 2048);
  hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(MLP_THREADS), 0, st, dx, n, L.wb, L.w, dout);
  return hipGetLastError();
}

// mean loss gradient of samples idx[0..n) (idx == nullptr: 0..n) into
// L.grad, mean loss into *dloss (device)
hipError_t learned_grad(LearnedState& L, hipStream_t st, const float* dx, const float* de,
                        const float* dtoi, const uint32_t* idx, int n, float* dloss) {
  if (n <= 0) return hipErrorInvalidValue;
  const int ntile = (n + MT - 1) / MT;
  CS_HIP_RET(ensure_train(L, ntile));
  hipLaunchKernelGGL(k_mlp_train_tile, dim3(ntile), dim3(MLP_THREADS), 0, st, dx, de, dtoi, idx,
                     n, L.wb, L.w, L.gpart, L.lpart);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((MF_NPARAM + 255) / 256), dim3(256), 0, st, L.gpart,
                     L.lpart, ntile, n, L.grad, dloss);
  return hipGetLastError();
}

hipError_t learned_adam_step(LearnedState& L, hipStream_t st, float lr) {
  L.adam_t++;
  const double c1 = 1.0 - std::pow(0.9, (double)L.adam_t);
  const double c2 = 1.0 - std::pow(0.999, (double)L.adam_t);
  hipLaunchKernelGGL(k_adam, dim3((MF_NPARAM + 255) / 256), dim3(256), 0, st, L.w, L.m, L.v,
                     L.grad, lr, (float)c1, (float)c2);
  hipLaunchKernelGGL(k_pack_weights, dim3((MB_TOTAL + 255) / 256), dim3(256), 0, st, L.w, L.wb);
  return hipGetLastError();
}

hipError_t learned_ensure_dataset(LearnedState& L, int64_t capacity) {
  if (capacity <= L.ds_alloc) return hipSuccess;
  lfree(L.dsx);
  lfree(L.dse);
  lfree(L.dst);
  L.ds_alloc = 0;
  CS_HIP_RET(lalloc(&L.dsx, (size_t)capacity * MF_IN));
  CS_HIP_RET(lalloc(&L.dse, (size_t)capacity));
  CS_HIP_RET(lalloc(&L.dst, (size_t)capacity));
  L.ds_alloc = capacity;
  return hipSuccess;
}

int64_t learned_dataset_size(LearnedState& L, hipStream_t st) {
  if (!L.ds_count || L.ds_cap <= 0) return 0;
  unsigned long long c = 0;
  if (hipMemcpyAsync(&c, L.ds_count, sizeof(c), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
  if (hipStreamSynchronize(st) != hipSuccess) return -1;
  return std::min<int64_t>((int64_t)c, L.ds_cap);
}

// Adam over seeded shuffles of the collected dataset; loss[e] = mean of the
// batch losses of epoch e.  The classes are balanced by repetition: every
// sample with a root appears r = round(#without / #with) times per epoch
// (1 <= r <= 64), so both labels weigh about the same in the loss.
hipError_t learned_train(LearnedState& L, hipStream_t st, int epochs, int batch, float lr,
                         uint64_t seed, float* host_loss) {
  const int64_t n64 = learned_dataset_size(L, st);
  if (n64 <= 0 || n64 > INT32_MAX / 4 || batch <= 0 || epochs < 0) return hipErrorInvalidValue;
  const int n = (int)n64;
  uint64_t *k64 = nullptr, *k64t = nullptr;
  uint32_t *order = nullptr, *ordt = nullptr, *k32 = nullptr, *k32t = nullptr, *bidx = nullptr,
           *bidxt = nullptr, *hist = nullptr, *part = nullptr, *ext = nullptr;
  float* bloss = nullptr;
  bool in_tmp = false;
  hipError_t e = hipSuccess;
  do {
    // canonical order of the samples (a hash of their contents)
    const size_t hw0 = radix_hist_words(n), pw0 = scan_partial_words((int)hw0);
    if ((e = lalloc(&k64, n)) || (e = lalloc(&k64t, n)) || (e = lalloc(&order, n)) ||
        (e = lalloc(&ordt, n)) || (e = lalloc(&hist, hw0)) || (e = lalloc(&part, pw0)))
      break;
    hipLaunchKernelGGL(k_ds_keys, dim3((n + 255) / 256), dim3(256), 0, st, L.dsx, L.dse, L.dst,
                       n, k64, order);
    if ((e = radix_sort_u64(st, k64, order, k64t, ordt, hist, part, n, 64, &in_tmp))) break;
    std::vector<uint32_t> ord((size_t)n);
    std::vector<float> lab((size_t)n);
    if ((e = hipMemcpyAsync(ord.data(), in_tmp ? ordt : order, sizeof(uint32_t) * n,
                            hipMemcpyDeviceToHost, st)))
      break;
    if ((e = hipMemcpyAsync(lab.data(), L.dse, sizeof(float) * n, hipMemcpyDeviceToHost, st))) break;
    if ((e = hipStreamSynchronize(st))) break;
    int64_t npos = 0;
    for (float v : lab) npos += v > 0.5f;
    const int64_t nneg = n - npos;
    const int r = npos > 0 ? (int)std::min<int64_t>(64, std::max<int64_t>(1, (nneg + npos / 2) / npos)) : 1;
    std::vector<uint32_t> ext_h;
    ext_h.reserve((size_t)(nneg + r * npos));
    for (uint32_t i : ord) {
      const int rep = lab[i] > 0.5f ? r : 1;
      for (int k = 0; k < rep; ++k) ext_h.push_back(i);
    }
    const int m = (int)ext_h.size();
    const int nbatch = (m + batch - 1) / batch;
    lfree(hist);
    lfree(part);
    const size_t hw = radix_hist_words(m), pw = scan_partial_words((int)hw);
    if ((e = lalloc(&ext, m)) || (e = lalloc(&k32, m)) || (e = lalloc(&k32t, m)) ||
        (e = lalloc(&bidx, m)) || (e = lalloc(&bidxt, m)) || (e = lalloc(&hist, hw)) ||
        (e = lalloc(&part, pw)) || (e = lalloc(&bloss, nbatch)))
      break;
    if ((e = hipMemcpyAsync(ext, ext_h.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice, st)))
      break;
    std::vector<float> bl((size_t)nbatch);
    for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) {
      hipLaunchKernelGGL(k_shuffle_keys, dim3((m + 255) / 256), dim3(256), 0, st, ext, m,
                         mix64(seed + 0x1000003ull * (uint64_t)(ep + 1)), k32, bidx);
      if ((e = radix_sort_u32(st, k32, bidx, k32t, bidxt, hist, part, m, 32, &in_tmp))) break;
      const uint32_t* bi = in_tmp ? bidxt : bidx;
      for (int b = 0; b < nbatch && e == hipSuccess; ++b) {
        const int bs = std::min(batch, m - b * batch);
        e = learned_grad(L, st, L.dsx, L.dse, L.dst, bi + (int64_t)b * batch, bs, bloss + b);
        if (e == hipSuccess) e = learned_adam_step(L, st, lr);
      }
      if (e) break;
      if ((e = hipMemcpyAsync(bl.data(), bloss, sizeof(float) * nbatch, hipMemcpyDeviceToHost, st)))
        break;
      if ((e = hipStreamSynchronize(st))) break;
      double sum = 0;
      for (float v : bl) sum += v;
      if (host_loss) host_loss[ep] = (float)(sum / nbatch);
    }
  } while (false);
  (void)hipStreamSynchronize(st);
  lfree(k64); lfree(k64t); lfree(order); lfree(ordt); lfree(k32); lfree(k32t);
  lfree(bidx); lfree(bidxt); lfree(hist); lfree(part); lfree(bloss); lfree(ext);
  return e;
}

}  // namespace cs
