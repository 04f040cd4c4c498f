// k_mlp.hip -- the learned smoothing-length predictor: a 6 -> 64 -> 64 -> 1
// tanh MLP with BF16 weights and FP32 accumulation on MFMA, its backward
// pass and Adam.
//
// MFMA forms (wave64): the K = 64 products use gfx950's
// v_mfma_f32_32x32x16_bf16, where lane l supplies A[l%32][8*(l/32)+i] and
// B[8*(l/32)+i][l%32], i = 0..7; layer 1 (K = 8) uses
// v_mfma_f32_32x32x8_bf16, lane l supplying A[l%32][4*(l/32)+i] and
// B[4*(l/32)+i][l%32], i = 0..3.  In both, result register r of lane l is
// D[8*(r/4) + 4*(l/32) + r%4][l%32].
//
// Forward: one wavefront per tile of 32 samples.  Layer 1 (K = 8, inputs
// padded from 6) and layer 2 (K = 64, four K = 16 products) are MFMA
// products; activations go through LDS to change from the result layout to
// the operand layout; the 64 -> 1 output layer is a VALU dot product.
// Backward: persistent workgroups of two wavefronts (64 samples) recompute
// the forward pass, then dW2 = H1^T dZ2 (K = 64 samples) and dH1 = dZ2 W2^T
// run on MFMA, the rest on VALU; partial gradients stay in registers/LDS
// across the workgroup's tiles and are added to the global gradient once.
#include "fsph_internal.h"

namespace fsph {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

constexpr int kIn = 6, kInP = 8, kH = 64;

__device__ __host__ inline unsigned short f2bf(float f) {
  unsigned u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short b) {
  unsigned u = (unsigned)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

__device__ inline f32x16 mfma(bf16x4 a, bf16x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}
__device__ inline f32x16 mfma16(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ inline __bf16 as_bf(unsigned short u) {
  return __builtin_bit_cast(__bf16, u);
}

__device__ inline int c_row(int r, int l) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }

}  // namespace

// flat FP32 parameter vector layout
constexpr int kOffW1 = 0;                      // [kInP][kH] (rows 6,7 zero)
constexpr int kOffB1 = kOffW1 + kInP * kH;     // [kH]
constexpr int kOffW2 = kOffB1 + kH;            // [kH][kH]  (k, n)
constexpr int kOffB2 = kOffW2 + kH * kH;       // [kH]
constexpr int kOffW3 = kOffB2 + kH;            // [kH]
constexpr int kOffB3 = kOffW3 + kH;            // [1]
constexpr int kNParams = kOffB3 + 1;

struct MlpW {
  const unsigned short *W1b;   // [kInP][kH] bf16
  const unsigned short *W2b;   // [kH][kH]   bf16 (k, n)
  const unsigned short *W2bT;  // [kH][kH]   bf16 (n, k)
  const float *p;              // FP32 master parameters (biases, w3, b3)
};

namespace {

// Layer 1 + 2 for the wave's 32 samples; leaves H1 (bf16) and H2 (fp32) in
// LDS, row-major [32][64].
__device__ inline void forward_tile(const MlpW &w, const float *X, int64_t B,
                                    int64_t p0, unsigned short *sH1,
                                    float *sH2) {
  const int l = threadIdx.x & 63;
  bf16x4 a;
  for (int i = 0; i < 4; ++i) {
    const int k = 4 * (l >> 5) + i;
    const int64_t p = p0 + (l & 31);
    a[i] = (short)f2bf((p < B && k < kIn) ? X[p * kInP + k] : 0.f);
  }
  for (int nt = 0; nt < 2; ++nt) {
    bf16x4 b;
    for (int i = 0; i < 4; ++i)
      b[i] = (short)w.W1b[(4 * (l >> 5) + i) * kH + nt * 32 + (l & 31)];
    f32x16 c = {0};
    c = mfma(a, b, c);
    const int col = nt * 32 + (l & 31);
    const float bias = w.p[kOffB1 + col];
    for (int r = 0; r < 16; ++r)
      sH1[c_row(r, l) * kH + col] = f2bf(tanhf(c[r] + bias));
  }
  __syncthreads();
  for (int nt = 0; nt < 2; ++nt) {
    f32x16 c = {0};
    for (int ks = 0; ks < kH / 16; ++ks) {
      bf16x8 aa, b;
      for (int i = 0; i < 8; ++i) {
        const int k = ks * 16 + 8 * (l >> 5) + i;
        aa[i] = as_bf(sH1[(l & 31) * kH + k]);
        b[i] = as_bf(w.W2b[k * kH + nt * 32 + (l & 31)]);
      }
      c = mfma16(aa, b, c);
    }
    const int col = nt * 32 + (l & 31);
    const float bias = w.p[kOffB2 + col];
    for (int r = 0; r < 16; ++r) sH2[c_row(r, l) * kH + col] = tanhf(c[r] + bias);
  }
  __syncthreads();
}

__device__ inline float output_dot(const MlpW &w, const float *h2row) {
  float y = w.p[kOffB3];
  for (int c = 0; c < kH; ++c) y += h2row[c] * w.p[kOffW3 + c];
  return y;
}

// One wavefront per 32 samples: y = MLP(x).  Every wave runs exactly one
// tile, so the block barriers inside are uniform.
__global__ __launch_bounds__(64) void k_mlp_fwd(MlpW w, const float *X,
                                                int64_t B, float *Y) {
  __shared__ unsigned short sH1[32 * kH];
  __shared__ float sH2[32 * kH];
  const int64_t p0 = (int64_t)blockIdx.x * 32;
  forward_tile(w, X, B, p0, sH1, sH2);
  const int l = threadIdx.x;
  if (l < 32 && p0 + l < B) Y[p0 + l] = output_dot(w, sH2 + l * kH);
}

// Gradient of L = mean (y - t)^2 over B samples, accumulated into g[] (FP32,
// zeroed by the caller); loss into loss[0] and mean t^2 (the loss of the
// zero predictor) into loss[1].  Workgroups of 128 threads
// (two wavefronts, 64 samples per round) loop over rounds.
__global__ __launch_bounds__(128) void k_mlp_bwd(MlpW w, const float *X,
                                                 const float *Tgt, int64_t B,
                                                 float *g, float *loss) {
  __shared__ unsigned short sH1[2][32 * kH];
  __shared__ float sH2[2][32 * kH];
  __shared__ unsigned short sdZ2[64 * kH];
  __shared__ float sdZ1[64 * kH];
  __shared__ float sX[64 * kInP];
  __shared__ float sdy[64];
  const int t = threadIdx.x, wv = t >> 6, l = t & 63;
  const int64_t rounds = (B + 63) / 64;
  // register partials: dW2 tile (wave wv owns output rows wv*32..+31)
  f32x16 acc[2] = {{0}, {0}};
  float gW1[3] = {0, 0, 0}, gb1 = 0, gb2 = 0, gw3 = 0, gb3 = 0, lsum = 0,
        tsum = 0;
  const float invB = 1.0f / (float)B;
  for (int64_t rd = blockIdx.x; rd < rounds; rd += gridDim.x) {
    const int64_t base = rd * 64;
    for (int q = t; q < 64 * kInP; q += 128) {
      const int64_t p = base + q / kInP;
      sX[q] = (p < B) ? X[base * kInP + q] : 0.f;
    }
    forward_tile(w, X, B, base + wv * 32, sH1[wv], sH2[wv]);
    // output layer and dy
    if (l < 32) {
      const int64_t p = base + wv * 32 + l;
      float dy = 0.f;
      if (p < B) {
        const float y = output_dot(w, sH2[wv] + l * kH);
        const float tp = Tgt[p];
        const float e = y - tp;
        lsum += e * e;
        tsum += tp * tp;
        dy = 2.f * e * invB;
      }
      sdy[wv * 32 + l] = dy;
    }
    __syncthreads();
    // dZ2 = dy w3 (1 - H2^2); dw3, db3, db2 partials (thread t < 64 owns
    // column t for the column sums)
    for (int q = t; q < 64 * kH; q += 128) {
      const int s = q / kH, cc = q % kH;
      const float h2 = sH2[s >> 5][(s & 31) * kH + cc];
      sdZ2[q] = f2bf(sdy[s] * w.p[kOffW3 + cc] * (1.f - h2 * h2));
    }
    if (t < kH) {
      for (int s = 0; s < 64; ++s) {
        const float h2 = sH2[s >> 5][(s & 31) * kH + t];
        gw3 += h2 * sdy[s];
        gb2 += sdy[s] * w.p[kOffW3 + t] * (1.f - h2 * h2);
      }
    }
    if (t == 0)
      for (int s = 0; s < 64; ++s) gb3 += sdy[s];
    __syncthreads();
    // dW2[k][n] += sum_s H1[s][k] dZ2[s][n]: wave wv owns k in wv*32..+31
    for (int nt = 0; nt < 2; ++nt) {
      for (int ks = 0; ks < 64 / 16; ++ks) {
        bf16x8 a, b;
        for (int i = 0; i < 8; ++i) {
          const int s = ks * 16 + 8 * (l >> 5) + i;
          a[i] = as_bf(sH1[s >> 5][(s & 31) * kH + wv * 32 + (l & 31)]);
          b[i] = as_bf(sdZ2[s * kH + nt * 32 + (l & 31)]);
        }
        acc[nt] = mfma16(a, b, acc[nt]);
      }
    }
    // dH1 = dZ2 W2^T for the wave's 32 samples; dZ1 = dH1 (1 - H1^2)
    for (int nt = 0; nt < 2; ++nt) {
      f32x16 c = {0};
      for (int ks = 0; ks < kH / 16; ++ks) {
        bf16x8 a, b;
        for (int i = 0; i < 8; ++i) {
          const int n = ks * 16 + 8 * (l >> 5) + i;
          a[i] = as_bf(sdZ2[(wv * 32 + (l & 31)) * kH + n]);
          b[i] = as_bf(w.W2bT[n * kH + nt * 32 + (l & 31)]);
        }
        c = mfma16(a, b, c);
      }
      const int col = nt * 32 + (l & 31);
      for (int r = 0; r < 16; ++r) {
        const int row = c_row(r, l);
        const float h1 = bf2f(sH1[wv][row * kH + col]);
        sdZ1[(wv * 32 + row) * kH + col] = c[r] * (1.f - h1 * h1);
      }
    }
    __syncthreads();
    // dW1[k][n] += sum_s X[s][k] dZ1[s][n]; db1[n] += sum_s dZ1[s][n]
    {
      const int n = t & 63, kh = t >> 6;  // kh: features 0-2 or 3-5
      for (int s = 0; s < 64; ++s) {
        const float d = sdZ1[s * kH + n];
        for (int j = 0; j < 3; ++j) gW1[j] += sX[s * kInP + kh * 3 + j] * d;
        if (kh == 0) gb1 += d;
      }
    }
    __syncthreads();
  }
  // flush partials
  for (int nt = 0; nt < 2; ++nt)
    for (int r = 0; r < 16; ++r) {
      const int k = wv * 32 + c_row(r, l);
      const int n = nt * 32 + (l & 31);
      atomicAdd(&g[kOffW2 + k * kH + n], acc[nt][r]);
    }
  {
    const int n = t & 63, kh = t >> 6;
    for (int j = 0; j < 3; ++j) atomicAdd(&g[kOffW1 + (kh * 3 + j) * kH + n], gW1[j]);
    if (kh == 0) atomicAdd(&g[kOffB1 + n], gb1);
  }
  if (t < kH) {
    atomicAdd(&g[kOffW3 + t], gw3);
    atomicAdd(&g[kOffB2 + t], gb2);
  }
  if (t == 0) atomicAdd(&g[kOffB3], gb3);
  // loss
  for (int o = 32; o > 0; o >>= 1) {
    lsum += __shfl_xor(lsum, o, 64);
    tsum += __shfl_xor(tsum, o, 64);
  }
  if (l == 0) {
    atomicAdd(loss, lsum * invB);
    atomicAdd(loss + 1, tsum * invB);
  }
}

// Adam (Kingma & Ba) on the FP32 master parameters, then the BF16 copies.
__global__ __launch_bounds__(256) void k_adam(float *p, float *m, float *v,
                                              const float *g, int n, float lr,
                                              float b1, float b2, float eps,
                                              float c1, float c2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  // the padding rows 6 and 7 of W1 stay zero
  if (i >= kOffW1 + kIn * kH && i < kOffB1) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void k_bf16_copies(const float *p,
                                                     unsigned short *W1b,
                                                     unsigned short *W2b,
                                                     unsigned short *W2bT) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < kInP * kH) W1b[i] = f2bf(p[kOffW1 + i]);
  if (i < kH * kH) {
    const int k = i / kH, n = i % kH;
    const unsigned short b = f2bf(p[kOffW2 + i]);
    W2b[i] = b;
    W2bT[n * kH + k] = b;
  }
}

}  // namespace

// ---- host side -------------------------------------------------------------
int mlp_ensure(Ctx *c) {
  auto &M = c->mlp;
  if (M.ready) return 0;
  FSPH_TRY(ensure(c, M.p, kNParams * 4));
  FSPH_TRY(ensure(c, M.m, kNParams * 4));
  FSPH_TRY(ensure(c, M.v, kNParams * 4));
  FSPH_TRY(ensure(c, M.g, (kNParams + 2) * 4));
  FSPH_TRY(ensure(c, M.W1b, kInP * kH * 2));
  FSPH_TRY(ensure(c, M.W2b, kH * kH * 2));
  FSPH_TRY(ensure(c, M.W2bT, kH * kH * 2));
  // seeded uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) initialisation
  std::vector<float> h(kNParams, 0.f);
  uint64_t s = 0x5EEDF1EC5ull ^ 0x9E3779B97F4A7C15ull;
  auto next = [&]() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int k = 0; k < kIn; ++k)
    for (int n = 0; n < kH; ++n)
      h[kOffW1 + k * kH + n] = (float)((2 * next() - 1) / std::sqrt((double)kIn));
  for (int i = 0; i < kH * kH; ++i)
    h[kOffW2 + i] = (float)((2 * next() - 1) / std::sqrt((double)kH));
  for (int i = 0; i < kH; ++i)
    h[kOffW3 + i] = (float)(0.1 * (2 * next() - 1) / std::sqrt((double)kH));
  FSPH_HIP(c, hipMemcpyAsync(M.p.p, h.data(), kNParams * 4,
                             hipMemcpyHostToDevice, c->stream));
  FSPH_HIP(c, hipMemsetAsync(M.m.p, 0, kNParams * 4, c->stream));
  FSPH_HIP(c, hipMemsetAsync(M.v.p, 0, kNParams * 4, c->stream));
  FSPH_HIP(c, hipStreamSynchronize(c->stream));
  M.step = 0;
  M.ready = true;
  return mlp_refresh(c);
}

int mlp_refresh(Ctx *c) {
  auto &M = c->mlp;
  k_bf16_copies<<<(kH * kH + 255) / 256, 256, 0, c->stream>>>(
      M.p.as<float>(), M.W1b.as<unsigned short>(), M.W2b.as<unsigned short>(),
      M.W2bT.as<unsigned short>());
  return hip_check(c, hipGetLastError(), "k_bf16_copies");
}

static MlpW weights_of(Ctx *c) {
  auto &M = c->mlp;
  return MlpW{M.W1b.as<unsigned short>(), M.W2b.as<unsigned short>(),
              M.W2bT.as<unsigned short>(), M.p.as<float>()};
}

int mlp_forward(Ctx *c, const float *X, int64_t B, float *Y) {
  FSPH_TRY(mlp_ensure(c));
  if (B <= 0) return 0;
  k_mlp_fwd<<<(unsigned)((B + 31) / 32), 64, 0, c->stream>>>(weights_of(c), X,
                                                             B, Y);
  return hip_check(c, hipGetLastError(), "k_mlp_fwd");
}

int mlp_gradient(Ctx *c, const float *X, const float *T, int64_t B) {
  FSPH_TRY(mlp_ensure(c));
  auto &M = c->mlp;
  FSPH_HIP(c, hipMemsetAsync(M.g.p, 0, (kNParams + 2) * 4, c->stream));
  if (B <= 0) return 0;
  const int64_t rounds = (B + 63) / 64;
  const unsigned grid = (unsigned)std::min<int64_t>(rounds, 512);
  k_mlp_bwd<<<grid, 128, 0, c->stream>>>(weights_of(c), X, T, B,
                                         M.g.as<float>(),
                                         M.g.as<float>() + kNParams);
  return hip_check(c, hipGetLastError(), "k_mlp_bwd");
}

int mlp_adam(Ctx *c, float lr) {
  auto &M = c->mlp;
  M.step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)M.step);
  const float c2 = 1.f - std::pow(b2, (float)M.step);
  k_adam<<<(kNParams + 255) / 256, 256, 0, c->stream>>>(
      M.p.as<float>(), M.m.as<float>(), M.v.as<float>(), M.g.as<float>(),
      kNParams, lr, b1, b2, 1e-8f, c1, c2);
  FSPH_HIP(c, hipGetLastError());
  return mlp_refresh(c);
}

int mlp_num_params() { return kNParams; }

void free_mlp(Ctx *c) {
  auto &M = c->mlp;
  DBuf *all[] = {&M.p, &M.m, &M.v, &M.g, &M.W1b, &M.W2b, &M.W2bT,
                 &M.X, &M.T, &M.Y, &M.idx, &M.norm};
  for (DBuf *b : all)
    if (b->p) {
      hipFree(b->p);
      b->p = nullptr;
      b->bytes = 0;
    }
}

}  // namespace fsph
