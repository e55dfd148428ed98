// mlp.hip -- the learned ruin proposal's network: a 2-layer MLP scored and
// trained on the GPU by its own kernels.
//
//   h = relu(W1 x + b1)      x: 16 fp32 features (zero-padded), W1: 256x16
//   y = w2 . h + b2          logit; trained with binary cross-entropy
//
// Weights are kept as fp32 masters (Adam state alongside) and used as bf16
// copies; the two 16x16x16 products of a 16-row batch tile (x W1^T in the
// forward pass, dH^T X for the W1 gradient) run on bf16 MFMA with fp32
// accumulation (v_mfma_f32_16x16x16bf16_1k).  Lane layout of that
// instruction (wave64): lane l holds A[l%16][4(l/16)+j] and
// B[4(l/16)+j][l%16] for j = 0..3, and returns D[4(l/16)+j][l%16].
// The second layer (256 -> 1) is a fp32 dot product and a wave reduction.
//
// One network per process (the engine's singleton convention); host
// buffers in and out, one stream owned here.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#define MLP_IN 16
#define MLP_H 256
#define MLP_P (MLP_H * MLP_IN + MLP_H + MLP_H + 1)   // W1, b1, w2, b2

#define CK(x)                                                               \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_),     \
              __FILE__, __LINE__);                                          \
      return -(100 + (int)e_);                                              \
    }                                                                       \
  } while (0)

typedef short v4i16 __attribute__((ext_vector_type(4)));
typedef float v4f32 __attribute__((ext_vector_type(4)));

// fp32 -> bf16 with round-to-nearest-even (finite inputs)
__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}

// forward: one wave per 16-row batch tile
__global__ void __launch_bounds__(64)
    k_mlp_fwd(const float *__restrict__ X, int B, const uint16_t *__restrict__ W1b,
              const float *__restrict__ P, const uint16_t *__restrict__ w2b,
              float *__restrict__ Hpre, float *__restrict__ Y) {
  const int lane = threadIdx.x;
  const int r0 = blockIdx.x * 16;
  const int li = lane & 15, kb = (lane >> 4) * 4;
  const float *b1 = P + MLP_H * MLP_IN;
  const float b2 = P[MLP_P - 1];
  v4i16 a;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int row = r0 + li;
    a[j] = (short)to_bf16(row < B ? X[(size_t)row * MLP_IN + kb + j] : 0.0f);
  }
  float yacc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int nt = 0; nt < MLP_H / 16; nt++) {
    const int n = nt * 16 + li;
    v4i16 b;
#pragma unroll
    for (int j = 0; j < 4; j++) b[j] = (short)W1b[n * MLP_IN + kb + j];
    v4f32 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    const float bn = b1[n], wn = from_bf16(w2b[n]);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const float h = c[j] + bn;
      const int row = r0 + kb + j;
      if (row < B) Hpre[(size_t)row * MLP_H + n] = h;
      yacc[j] += (h > 0.f ? h : 0.f) * wn;
    }
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    float v = yacc[j];
    for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
    yacc[j] = v;
  }
  if (li == 0)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int row = r0 + kb + j;
      if (row < B) Y[row] = yacc[j] + b2;
    }
}

// dy = (sigmoid(y) - t) / B, the mean binary cross-entropy, and db2
__global__ void __launch_bounds__(256)
    k_mlp_dy(const float *__restrict__ Y, const float *__restrict__ T, int B,
             float *__restrict__ dY, float *__restrict__ G, float *__restrict__ loss) {
  __shared__ float s1[256], s2[256];
  float g = 0.f, l = 0.f;
  for (int i = threadIdx.x; i < B; i += 256) {
    const float y = Y[i], t = T[i];
    const float p = 1.f / (1.f + expf(-y));
    const float d = (p - t) / (float)B;
    dY[i] = d;
    g += d;
    // log(1 + e^-|y|) + max(y, 0) - t y  (stable BCE with logits)
    l += (log1pf(expf(-fabsf(y))) + fmaxf(y, 0.f) - t * y) / (float)B;
  }
  s1[threadIdx.x] = g;
  s2[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] += s2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    G[MLP_P - 1] = s1[0];
    loss[0] = s2[0];
  }
}

// gradients of W1 (MFMA over the batch), b1 and w2: one wave per 16 hidden
// units, looping over the batch in tiles of 16 rows
__global__ void __launch_bounds__(64)
    k_mlp_bwd(const float *__restrict__ X, int B, const float *__restrict__ Hpre,
              const float *__restrict__ dY, const uint16_t *__restrict__ w2b,
              float *__restrict__ G) {
  const int lane = threadIdx.x;
  const int n0 = blockIdx.x * 16;
  const int li = lane & 15, kb = (lane >> 4) * 4;
  const float wn = from_bf16(w2b[n0 + li]);       // hidden unit of A's row li
  v4f32 c = {0.f, 0.f, 0.f, 0.f};
  float db1 = 0.f, dw2 = 0.f;
  for (int b0 = 0; b0 < B; b0 += 16) {
    v4i16 a, b;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int row = b0 + kb + j;
      float dh = 0.f, xv = 0.f;
      if (row < B) {
        const float hp = Hpre[(size_t)row * MLP_H + n0 + li];
        const float dy = dY[row];
        dh = hp > 0.f ? dy * wn : 0.f;
        db1 += dh;
        dw2 += dy * (hp > 0.f ? hp : 0.f);
        xv = X[(size_t)row * MLP_IN + li];
      }
      a[j] = (short)to_bf16(dh);     // A[m = li][kk = kb + j] = dH[row][n0 + li]
      b[j] = (short)to_bf16(xv);     // B[kk = kb + j][k = li] = X[row][li]
    }
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
  // D[m = kb + j][k = li] = dW1[n0 + kb + j][li]
#pragma unroll
  for (int j = 0; j < 4; j++) G[(n0 + kb + j) * MLP_IN + li] = c[j];
  db1 += __shfl_xor(db1, 16, 64);
  db1 += __shfl_xor(db1, 32, 64);
  dw2 += __shfl_xor(dw2, 16, 64);
  dw2 += __shfl_xor(dw2, 32, 64);
  if (lane < 16) {
    G[MLP_H * MLP_IN + n0 + li] = db1;
    G[MLP_H * MLP_IN + MLP_H + n0 + li] = dw2;
  }
}

// Adam on the fp32 masters, then the bf16 copies
__global__ void k_mlp_adam(float *__restrict__ P, const float *__restrict__ G,
                           float *__restrict__ M, float *__restrict__ V,
                           uint16_t *__restrict__ W1b, uint16_t *__restrict__ w2b,
                           float lr, float b1, float b2, float eps, float c1,
                           float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_P) return;
  const float g = G[i];
  const float m = b1 * M[i] + (1.f - b1) * g;
  const float v = b2 * V[i] + (1.f - b2) * g * g;
  M[i] = m;
  V[i] = v;
  const float p = P[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
  P[i] = p;
  if (i < MLP_H * MLP_IN) W1b[i] = to_bf16(p);
  else if (i >= MLP_H * MLP_IN + MLP_H && i < MLP_P - 1)
    w2b[i - MLP_H * MLP_IN - MLP_H] = to_bf16(p);
}

__global__ void k_mlp_round(const float *__restrict__ P, uint16_t *__restrict__ W1b,
                            uint16_t *__restrict__ w2b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < MLP_H * MLP_IN) W1b[i] = to_bf16(P[i]);
  if (i < MLP_H) w2b[i] = to_bf16(P[MLP_H * MLP_IN + MLP_H + i]);
}

namespace {
struct Net {
  bool live = false;
  int maxb = 0;
  hipStream_t st = nullptr;
  float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;
  uint16_t *W1b = nullptr, *w2b = nullptr;
  float *X = nullptr, *T = nullptr, *Y = nullptr, *dY = nullptr;
  float *H = nullptr, *loss = nullptr;
  float *hbuf = nullptr;   // pinned staging
  long long steps = 0;
};
Net N;

void net_release() {
  if (N.st) (void)hipStreamSynchronize(N.st);
  void *bufs[] = {N.P, N.G, N.M, N.V, N.W1b, N.w2b, N.X, N.T, N.Y, N.dY, N.H, N.loss};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (N.hbuf) (void)hipHostFree(N.hbuf);
  if (N.st) (void)hipStreamDestroy(N.st);
  N = Net();
}

int put(float *dst, const float *src, size_t n) {
  memcpy(N.hbuf, src, n * 4);
  CK(hipMemcpyAsync(dst, N.hbuf, n * 4, hipMemcpyHostToDevice, N.st));
  CK(hipStreamSynchronize(N.st));
  return 0;
}

int get(float *dst, const float *src, size_t n) {
  CK(hipMemcpyAsync(N.hbuf, src, n * 4, hipMemcpyDeviceToHost, N.st));
  CK(hipStreamSynchronize(N.st));
  memcpy(dst, N.hbuf, n * 4);
  return 0;
}

int forward(const float *X, int B) {
  if (!N.live || B < 1 || B > N.maxb || !X) return -1;
  if (int rc = put(N.X, X, (size_t)B * MLP_IN)) return rc;
  hipLaunchKernelGGL(k_mlp_fwd, dim3((B + 15) / 16), dim3(64), 0, N.st, N.X, B,
                     N.W1b, N.P, N.w2b, N.H, N.Y);
  CK(hipGetLastError());
  return 0;
}
}  // namespace

#pragma GCC visibility push(default)
extern "C" {

// Allocate the network for batches of up to max_batch rows (weights zero
// until cc_mlp_set_params).
int cc_mlp_init(int max_batch) {
  if (max_batch < 1 || max_batch > (1 << 20)) return -1;
  if (N.live) net_release();
  N.maxb = max_batch;
  CK(hipStreamCreateWithFlags(&N.st, hipStreamNonBlocking));
  const size_t pb = (size_t)MLP_P * 4;
  CK(hipMalloc((void **)&N.P, pb));
  CK(hipMalloc((void **)&N.G, pb));
  CK(hipMalloc((void **)&N.M, pb));
  CK(hipMalloc((void **)&N.V, pb));
  CK(hipMemset(N.P, 0, pb));
  CK(hipMemset(N.G, 0, pb));
  CK(hipMemset(N.M, 0, pb));
  CK(hipMemset(N.V, 0, pb));
  CK(hipMalloc((void **)&N.W1b, MLP_H * MLP_IN * 2));
  CK(hipMalloc((void **)&N.w2b, MLP_H * 2));
  CK(hipMemset(N.W1b, 0, MLP_H * MLP_IN * 2));
  CK(hipMemset(N.w2b, 0, MLP_H * 2));
  CK(hipMalloc((void **)&N.X, (size_t)max_batch * MLP_IN * 4));
  CK(hipMalloc((void **)&N.T, (size_t)max_batch * 4));
  CK(hipMalloc((void **)&N.Y, (size_t)max_batch * 4));
  CK(hipMalloc((void **)&N.dY, (size_t)max_batch * 4));
  CK(hipMalloc((void **)&N.H, (size_t)max_batch * MLP_H * 4));
  CK(hipMalloc((void **)&N.loss, 64));
  const size_t hb = (size_t)max_batch * MLP_IN * 4 > pb ? (size_t)max_batch * MLP_IN * 4 : pb;
  CK(hipHostMalloc((void **)&N.hbuf, hb, hipHostMallocDefault));
  N.steps = 0;
  N.live = true;
  return 0;
}

int cc_mlp_free(void) {
  if (N.live) net_release();
  return 0;
}

int cc_mlp_param_count(void) { return MLP_P; }

// p: MLP_P floats, W1 (256 x 16 row-major), b1 (256), w2 (256), b2; resets
// the optimizer state
int cc_mlp_set_params(const float *p) {
  if (!N.live || !p) return -1;
  if (int rc = put(N.P, p, MLP_P)) return rc;
  CK(hipMemsetAsync(N.M, 0, (size_t)MLP_P * 4, N.st));
  CK(hipMemsetAsync(N.V, 0, (size_t)MLP_P * 4, N.st));
  hipLaunchKernelGGL(k_mlp_round, dim3((MLP_H * MLP_IN + 255) / 256), dim3(256),
                     0, N.st, N.P, N.W1b, N.w2b);
  CK(hipGetLastError());
  CK(hipStreamSynchronize(N.st));
  N.steps = 0;
  return 0;
}

int cc_mlp_get_params(float *p) {
  if (!N.live || !p) return -1;
  return get(p, N.P, MLP_P);
}

// logits of B feature rows (X: B x 16 fp32)
int cc_mlp_forward(const float *X, int B, float *y) {
  if (!y) return -1;
  if (int rc = forward(X, B)) return rc;
  return get(y, N.Y, (size_t)B);
}

// gradient of the mean BCE over B rows with targets t (0/1); g (optional)
// receives MLP_P floats in the parameter layout, loss (optional) the loss
int cc_mlp_grad(const float *X, const float *t, int B, float *g, float *loss) {
  if (!t) return -1;
  if (int rc = forward(X, B)) return rc;
  if (int rc = put(N.T, t, (size_t)B)) return rc;
  hipLaunchKernelGGL(k_mlp_dy, dim3(1), dim3(256), 0, N.st, N.Y, N.T, B, N.dY,
                     N.G, N.loss);
  CK(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_bwd, dim3(MLP_H / 16), dim3(64), 0, N.st, N.X, B,
                     N.H, N.dY, N.w2b, N.G);
  CK(hipGetLastError());
  if (g)
    if (int rc = get(g, N.G, MLP_P)) return rc;
  if (loss)
    if (int rc = get(loss, N.loss, 1)) return rc;
  CK(hipStreamSynchronize(N.st));
  return 0;
}

// one Adam step from the last cc_mlp_grad
int cc_mlp_adam(float lr, float beta1, float beta2, float eps) {
  if (!N.live) return -1;
  N.steps++;
  const float c1 = 1.f - powf(beta1, (float)N.steps);
  const float c2 = 1.f - powf(beta2, (float)N.steps);
  hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_P + 255) / 256), dim3(256), 0, N.st,
                     N.P, N.G, N.M, N.V, N.W1b, N.w2b, lr, beta1, beta2, eps, c1,
                     c2);
  CK(hipGetLastError());
  CK(hipStreamSynchronize(N.st));
  return 0;
}

}  // extern "C"
#pragma GCC visibility pop
