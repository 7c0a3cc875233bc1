// Warm-start network of the local return map: an MLP 18 -> 64 -> 64 -> 64 -> 12 with SiLU,
// BF16 weights and activations on the matrix cores with FP32 accumulation (no counterpart in the
// reference, whose local loop starts from zero slip, calculatePlasticity.cc:311).
//
// Inputs: the 12 trial ratios (tau_a - W_a)/s_a and the lattice-frame increment of the elastic
// Green-Lagrange strain (6, Voigt 11,22,33,23,13,12), normalised as (x - mean) * scale.
// Outputs: 12 slip increments in units of out_scale.
//
// Every layer runs transposed, Z^T = W^T X^T, one wave per 16 samples, on
// v_mfma_f32_16x16x16_bf16 (A: 16x16 weight tile, B: 16 features x 16 samples).  Lane l of the
// result holds Z[feature 4(l/16)+j][sample l%16], j = 0..3, which is exactly the B-operand layout
// of the next layer's K-chunk, so activations stay in registers between layers.  The weight
// gradients need the samples on the reduction axis; each wave stages its activations and output
// gradients through LDS as [feature][16 samples] and reads them back in A/B layouts.
//
// Parameter blob (FP32 master copy): per layer l, W^T[N_l][K_l] row-major then b[N_l], with the
// padded sizes (K, N) = (32, 64), (64, 64), (64, 64), (64, 16); then in_mean[18], in_scale[18],
// out_scale[12].  Padded weights are zero and stay zero (their gradients are zero).
#pragma once
#include <hip/hip_runtime.h>

namespace mlp {

constexpr int NIN = 18, NH = 64, NOUT = 12;
constexpr int LK[4] = {32, 64, 64, 64};
constexpr int LN[4] = {64, 64, 64, 16};
constexpr int OFF_W[4] = {0, 2112, 6272, 10432};
constexpr int OFF_B[4] = {2048, 6208, 10368, 11456};
constexpr int NPARAM = 11472;
constexpr int OFF_MEAN = NPARAM, OFF_SCALE = NPARAM + NIN, OFF_OSCALE = NPARAM + 2 * NIN;
constexpr int NBLOB = NPARAM + 2 * NIN + NOUT;  // 11520
constexpr int NWBF = 2048 + 4096 + 4096 + 1024;  // bf16 weights of one layout

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (inputs are finite)
__host__ __device__ inline unsigned short f2bf(float f) {
  unsigned u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

__device__ inline v4s pack4(float a, float b, float c, float d) {
  v4s r;
  r[0] = (short)f2bf(a); r[1] = (short)f2bf(b); r[2] = (short)f2bf(c); r[3] = (short)f2bf(d);
  return r;
}
__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ inline float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s + z * s * (1.0f - s);
}
__device__ inline v4f mfma(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// A-operand fragment of tile (row tile rt, K chunk kc) of a row-major bf16 matrix with leading
// dimension ld: lane l reads M[16 rt + l%16][16 kc + 4(l/16) .. +3] (one 8-byte load).
__device__ inline v4s afrag(const unsigned short *M, int ld, int rt, int kc, int lane) {
  return *reinterpret_cast<const v4s *>(M + (16 * rt + (lane & 15)) * ld + 16 * kc + 4 * (lane >> 4));
}

// Device weights: bf16 W^T (out-major, forward) and W (in-major, backward) of each layer, packed
// layer after layer in the same order as the blob; plus the FP32 blob (biases, normalisation).
struct Weights {
  const unsigned short *wt;  // NWBF
  const unsigned short *w;   // NWBF
  const float *blob;         // NBLOB
};
__host__ __device__ constexpr int wbf_off(int l) { return l == 0 ? 0 : l == 1 ? 2048 : l == 2 ? 6144 : 10240; }

// Input fragment of a wave: lane l holds normalised feature 16c + 4(l/16) + j of sample l%16.
__device__ inline void load_input(const float *__restrict__ X, long long n, long long s0, const float *blob, int lane,
                                  v4s act0[2]) {
  const long long s = s0 + (lane & 15);
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int f = 16 * c + 4 * (lane >> 4) + j;
      v[j] = (f < NIN && s < n) ? (X[s * NIN + f] - blob[OFF_MEAN + f]) * blob[OFF_SCALE + f] : 0.0f;
    }
    act0[c] = pack4(v[0], v[1], v[2], v[3]);
  }
}

// One hidden layer: z = W^T act + b (NT output tiles, KC input chunks); act_out = bf16(silu(z)).
template <int L, int NT, int KC>
__device__ inline void hidden(const Weights &W, const v4s *act_in, v4f *z, v4s *act_out, int lane) {
  const unsigned short *wt = W.wt + wbf_off(L);
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    v4f acc;
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = W.blob[OFF_B[L] + 16 * t + 4 * (lane >> 4) + j];
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) acc = mfma(afrag(wt, LK[L], t, kc, lane), act_in[kc], acc);
    z[t] = acc;
    act_out[t] = pack4(silu(acc[0]), silu(acc[1]), silu(acc[2]), silu(acc[3]));
  }
}

// Forward pass of 16 samples; keeps what the backward pass needs.
struct Fwd {
  v4s a0[2], a1[4], a2[4], a3[4];
  v4f z1[4], z2[4], z3[4], y;
};
__device__ inline void forward(const Weights &W, Fwd &f, int lane) {
  hidden<0, 4, 2>(W, f.a0, f.z1, f.a1, lane);
  hidden<1, 4, 4>(W, f.a1, f.z2, f.a2, lane);
  hidden<2, 4, 4>(W, f.a2, f.z3, f.a3, lane);
  const unsigned short *wt = W.wt + wbf_off(3);
  v4f acc;
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = W.blob[OFF_B[3] + 4 * (lane >> 4) + j];
#pragma unroll
  for (int kc = 0; kc < 4; ++kc) acc = mfma(afrag(wt, 64, 0, kc, lane), f.a3[kc], acc);
  f.y = acc;
}

// Inference: x0[s][o] = y[o][s] * out_scale[o] for o < 12.  One wave per 16-sample tile, grid-stride.
__global__ void __launch_bounds__(256) k_mlp_predict(Weights W, long long n, const float *__restrict__ X, float *__restrict__ Y) {
  const int lane = threadIdx.x & 63;
  const long long wave = (blockIdx.x * (long long)blockDim.x + threadIdx.x) >> 6;
  const long long nwave = ((long long)gridDim.x * blockDim.x) >> 6;
  const long long ntile = (n + 15) / 16;
  for (long long t = wave; t < ntile; t += nwave) {
    Fwd f;
    load_input(X, n, 16 * t, W.blob, lane, f.a0);
    forward(W, f, lane);
    const long long s = 16 * t + (lane & 15);
    if (s < n) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int o = 4 * (lane >> 4) + j;
        if (o < NOUT) Y[s * NOUT + o] = f.y[j] * W.blob[OFF_OSCALE + o];
      }
    }
  }
}

// sum over the 16 lanes of a row group (lanes with equal l/16)
__device__ inline float rowsum16(float v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 8);
  return v;
}

// Stage a C-layout tile set (NT tiles of 16 features x 16 samples) into LDS as [feature][sample] bf16.
template <int NT>
__device__ inline void stage(unsigned short *lds, const v4s *v, int lane) {
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) lds[(16 * t + 4 * (lane >> 4) + j) * 16 + (lane & 15)] = (unsigned short)v[t][j];
}

// dW^T tiles of layer L: dWT[16 rt + 4(l/16)+j][16 kc + l%16] = sum_s G[out][s] A[in][s], from the
// staged G (NT x 16 features) and A (KC x 16 features); bias gradients from the FP32 G.
template <int L, int NT, int KC>
__device__ inline void weight_grad(const unsigned short *lg, const unsigned short *la, const v4f *G, float *gout, int lane) {
#pragma unroll
  for (int rt = 0; rt < NT; ++rt) {
    const v4s a = *reinterpret_cast<const v4s *>(lg + (16 * rt + (lane & 15)) * 16 + 4 * (lane >> 4));
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      const v4s b = *reinterpret_cast<const v4s *>(la + (16 * kc + (lane & 15)) * 16 + 4 * (lane >> 4));
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma(a, b, acc);
#pragma unroll
      for (int j = 0; j < 4; ++j) gout[OFF_W[L] + (16 * rt + 4 * (lane >> 4) + j) * LK[L] + 16 * kc + (lane & 15)] = acc[j];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float bs = rowsum16(G[rt][j]);
      if ((lane & 15) == 0) gout[OFF_B[L] + 16 * rt + 4 * (lane >> 4) + j] = bs;
    }
  }
}

// Back-propagation through layer L: dA^T[in][s] = sum_out W[in][out] G[out][s] (KT input tiles,
// NC output chunks), then G_in = dA * silu'(z).
template <int L, int KT, int NC>
__device__ inline void back_act(const Weights &W, const v4s *Gb, const v4f *z, v4f *Gin, int lane) {
  const unsigned short *w = W.w + wbf_off(L);
#pragma unroll
  for (int t = 0; t < KT; ++t) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < NC; ++kc) acc = mfma(afrag(w, LN[L], t, kc, lane), Gb[kc], acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) Gin[t][j] = acc[j] * dsilu(z[t][j]);
  }
}

// Forward + backward of one 16-sample tile per wave (4 waves per block).  Loss: sum over valid
// samples and the 12 outputs of (y - t)^2 in normalised units, times inv_count.  Each wave writes
// its parameter gradient (NPARAM floats) to part[wave] and its loss to lpart[wave].
// Samples are X[idx(k)], idx(k) = (mul * (base + k) + add) mod n for k < count (a permutation of
// the data set when gcd(mul, n) = 1).
__global__ void __launch_bounds__(256) k_mlp_grad(Weights W, long long n, const float *__restrict__ X,
                                                  const float *__restrict__ Yt, long long base, long long count,
                                                  long long mul, long long add, float inv_count, float *__restrict__ part,
                                                  float *__restrict__ lpart) {
  __shared__ unsigned short lds[4][2][64 * 16];
  const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
  const long long wave = (blockIdx.x * (long long)blockDim.x + threadIdx.x) >> 6;
  unsigned short *lg = lds[wib][0], *la = lds[wib][1];
  float *gout = part + wave * NPARAM;
  const long long k = 16 * wave + (lane & 15);
  const bool valid = k < count;
  const long long s = valid ? (long long)(((unsigned long long)mul * (unsigned long long)(base + k) + (unsigned long long)add) % (unsigned long long)n) : 0;

  Fwd f;
  // input fragment of this lane's sample (gathered)
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int fi = 16 * c + 4 * (lane >> 4) + j;
      v[j] = (fi < NIN && valid) ? (X[s * NIN + fi] - W.blob[OFF_MEAN + fi]) * W.blob[OFF_SCALE + fi] : 0.0f;
    }
    f.a0[c] = pack4(v[0], v[1], v[2], v[3]);
  }
  forward(W, f, lane);

  // output gradient
  v4f G4[1];
  float lsum = 0.f;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int o = 4 * (lane >> 4) + j;
    float g = 0.f;
    if (valid && o < NOUT) {
      const float d = f.y[j] - Yt[s * NOUT + o] / W.blob[OFF_OSCALE + o];
      lsum += d * d;
      g = 2.0f * d * inv_count;
    }
    G4[0][j] = g;
  }
  // wave loss
  for (int m = 32; m >= 1; m >>= 1) lsum += __shfl_xor(lsum, m);
  if (lane == 0) lpart[wave] = lsum * inv_count;

  v4s Gb4[1] = {pack4(G4[0][0], G4[0][1], G4[0][2], G4[0][3])};
  // layer 4 (index 3): dW from G4 and a3
  stage<1>(lg, Gb4, lane);
  stage<4>(la, f.a3, lane);
  __syncthreads();
  weight_grad<3, 1, 4>(lg, la, G4, gout, lane);
  __syncthreads();
  v4f G3[4];
  back_act<3, 4, 1>(W, Gb4, f.z3, G3, lane);
  v4s Gb3[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) Gb3[t] = pack4(G3[t][0], G3[t][1], G3[t][2], G3[t][3]);
  stage<4>(lg, Gb3, lane);
  stage<4>(la, f.a2, lane);
  __syncthreads();
  weight_grad<2, 4, 4>(lg, la, G3, gout, lane);
  __syncthreads();
  v4f G2[4];
  back_act<2, 4, 4>(W, Gb3, f.z2, G2, lane);
  v4s Gb2[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) Gb2[t] = pack4(G2[t][0], G2[t][1], G2[t][2], G2[t][3]);
  stage<4>(lg, Gb2, lane);
  stage<4>(la, f.a1, lane);
  __syncthreads();
  weight_grad<1, 4, 4>(lg, la, G2, gout, lane);
  __syncthreads();
  v4f G1[4];
  back_act<1, 4, 4>(W, Gb2, f.z1, G1, lane);
  v4s Gb1[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) Gb1[t] = pack4(G1[t][0], G1[t][1], G1[t][2], G1[t][3]);
  stage<4>(lg, Gb1, lane);
  stage<2>(la, f.a0, lane);
  __syncthreads();
  weight_grad<0, 4, 2>(lg, la, G1, gout, lane);
}

// Sum of the per-wave gradients (and losses) in wave order; mode 0 stores the gradient, mode 1
// applies one Adam step to the master weights (lr_t includes the bias corrections) and refreshes
// both bf16 layouts.  One thread per parameter.
__global__ void __launch_bounds__(256) k_mlp_reduce(int mode, int nwave, const float *__restrict__ part,
                                                    const float *__restrict__ lpart, float *grad, float *blob, float *m1,
                                                    float *m2, unsigned short *wt, unsigned short *w, float lr_t,
                                                    float beta1, float beta2, float eps, double *loss_acc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0 && loss_acc) {
    double l = 0.0;
    for (int i = 0; i < nwave; ++i) l += lpart[i];
    *loss_acc += l;
  }
  if (p >= NPARAM) return;
  float g = 0.f;
  for (int i = 0; i < nwave; ++i) g += part[(long long)i * NPARAM + p];
  if (mode == 0) { grad[p] = g; return; }
  const float a = beta1 * m1[p] + (1.f - beta1) * g;
  const float b = beta2 * m2[p] + (1.f - beta2) * g * g;
  m1[p] = a;
  m2[p] = b;
  const float v = blob[p] - lr_t * a / (sqrtf(b) + eps);
  blob[p] = v;
  // bf16 copies of weights (biases stay FP32)
  int L = 3;
  while (L > 0 && p < OFF_W[L]) --L;
  if (p < OFF_B[L]) {
    const int r = (p - OFF_W[L]) / LK[L], c = (p - OFF_W[L]) % LK[L];
    const unsigned short h = f2bf(v);
    wt[wbf_off(L) + r * LK[L] + c] = h;
    w[wbf_off(L) + c * LN[L] + r] = h;
  }
}

// Refresh both bf16 layouts from the master blob.
__global__ void k_mlp_cast(const float *__restrict__ blob, unsigned short *wt, unsigned short *w) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  int L = 3;
  while (L > 0 && p < OFF_W[L]) --L;
  if (p < OFF_B[L]) {
    const int r = (p - OFF_W[L]) / LK[L], c = (p - OFF_W[L]) % LK[L];
    const unsigned short h = f2bf(blob[p]);
    wt[wbf_off(L) + r * LK[L] + c] = h;
    w[wbf_off(L) + c * LN[L] + r] = h;
  }
}

}  // namespace mlp
