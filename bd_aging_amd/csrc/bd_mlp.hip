// bd_mlp.hip -- the learned surrogate of F_s(k, t_w, t): an ensemble of MLPs
// n_in -> hidden -> hidden -> n_out with GELU, trained and evaluated by
// hand-written gfx950 kernels.  No reference counterpart.
//
// Numerics: BF16 weights and activations feed the matrix cores
// (v_mfma_f32_32x32x16_bf16) with FP32 accumulation; pre-activations, the
// output layer, gradients, Adam moments and the master weights are FP32.
// All members are processed by the same launches (blockIdx.z = member).
//
// Parameter packing per member (FP32 master copy, row-major, x @ W):
//   W1[n_in][H] b1[H] W2[H][H] b2[H] W3[H][n_out] b3[n_out]
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bd_abi.h"
#include "bd_internal.h"

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef short s16x8_t __attribute__((ext_vector_type(8)));
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16_t;

struct bd_mlp {
  int members = 0, n_in = 0, hidden = 0, n_out = 0;
  size_t nparam = 0;  // per member
  size_t off_w1 = 0, off_b1 = 0, off_w2 = 0, off_b2 = 0, off_w3 = 0, off_b3 = 0;
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;  // members * nparam
  unsigned short* wb = nullptr;  // bf16 copies of W1, W2, W3 (members * nparam layout)
  int adam_t = 0;
  int cap = 0;  // batch capacity of the activation buffers
  unsigned short *xb = nullptr, *h1b = nullptr, *h2b = nullptr, *dyb = nullptr,
                 *dz2b = nullptr, *dz1b = nullptr;
  float *x = nullptr, *t = nullptr, *z1 = nullptr, *z2 = nullptr, *y = nullptr, *dy = nullptr,
        *dh = nullptr, *dz = nullptr, *dz1 = nullptr, *loss = nullptr;
  // device-resident data set
  float *dx = nullptr, *dt = nullptr;
  int n_data = 0;
  int* idx = nullptr;
  float* lhist = nullptr;  // per-batch losses of one training epoch
  hipStream_t stream = nullptr;
};

namespace {

constexpr int kEw = 256;

#define MLP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return bdi::set_error(BD_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

__host__ __device__ inline unsigned short f2bf(float f) {
  // round to nearest even (finite inputs)
  unsigned int u;
  __builtin_memcpy(&u, &f, 4);
  const unsigned int r = ((u >> 16) & 1u) + 0x7fffu;
  return (unsigned short)((u + r) >> 16);
}

__device__ inline float bf2f(unsigned short b) {
  const unsigned int u = (unsigned int)b << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// C[m] (M x N, FP32) = A[m] (M x K) * B[m] (K x N), BF16 operands, arbitrary
// element strides.  One wavefront per 32 x 32 output tile; the K loop feeds
// v_mfma_f32_32x32x16_bf16: lane l holds A[row l%32][k0 + 8(l/32) + 0..7]
// and B[k0 + 8(l/32) + 0..7][col l%32]; accumulator element v of lane l is
// C[8(v/4) + 4(l/32) + v%4][l%32].
struct GemmArgs {
  const unsigned short* A;
  long sAm, sAi, sAk;
  const unsigned short* B;
  long sBm, sBk, sBj;
  float* C;
  long sCm, sCi, sCj;
  int M, N, K;
};

__global__ void __launch_bounds__(64) k_gemm_bf16(GemmArgs g) {
  const int m = (int)blockIdx.z;
  const int lane = (int)threadIdx.x;
  const int row = (int)blockIdx.y * 32 + lane % 32;
  const int col = (int)blockIdx.x * 32 + lane % 32;
  const int kh = 8 * (lane / 32);
  const unsigned short* A = g.A + (long)m * g.sAm;
  const unsigned short* B = g.B + (long)m * g.sBm;
  floatx16_t acc = {};
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    s16x8_t a, b;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int k = k0 + kh + i;
      a[i] = (row < g.M && k < g.K) ? (short)A[(long)row * g.sAi + (long)k * g.sAk] : (short)0;
      b[i] = (col < g.N && k < g.K) ? (short)B[(long)k * g.sBk + (long)col * g.sBj] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                  __builtin_bit_cast(bf16x8_t, b), acc, 0, 0, 0);
  }
  float* C = g.C + (long)m * g.sCm;
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int r = (int)blockIdx.y * 32 + 8 * (v / 4) + 4 * (lane / 32) + v % 4;
    if (r < g.M && col < g.N) C[(long)r * g.sCi + (long)col * g.sCj] = acc[v];
  }
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// z = c + b (per column), h = bf16(gelu(z)); c and z may alias.
__global__ void k_bias_gelu(const float* c, const float* p, size_t bias_off, size_t nparam,
                            int rows, int cols, float* z, unsigned short* h) {
  const int m = (int)blockIdx.y;
  const size_t n = (size_t)rows * cols;
  const size_t base = (size_t)m * n;
  const float* b = p + (size_t)m * nparam + bias_off;
  for (size_t k = blockIdx.x * (size_t)kEw + threadIdx.x; k < n; k += (size_t)gridDim.x * kEw) {
    const float v = c[base + k] + b[k % cols];
    z[base + k] = v;
    h[base + k] = f2bf(gelu(v));
  }
}

__global__ void k_bias(float* y, const float* p, size_t bias_off, size_t nparam, int rows,
                       int cols) {
  const int m = (int)blockIdx.y;
  const size_t n = (size_t)rows * cols;
  const float* b = p + (size_t)m * nparam + bias_off;
  for (size_t k = blockIdx.x * (size_t)kEw + threadIdx.x; k < n; k += (size_t)gridDim.x * kEw)
    y[(size_t)m * n + k] += b[k % cols];
}

// mean squared error per member; dy = 2 (y - t) / (rows * cols)
__global__ void k_loss_grad(const float* y, const float* t, int t_per_member, int rows, int cols,
                            float* dy, unsigned short* dyb, float* loss) {
  const int m = (int)blockIdx.x;
  const size_t n = (size_t)rows * cols;
  const float* tm = t + (t_per_member ? (size_t)m * n : 0);
  const float scale = 2.0f / (float)n;
  float acc = 0.0f;
  for (size_t k = threadIdx.x; k < n; k += kEw) {
    const float d = y[(size_t)m * n + k] - tm[k];
    acc += d * d;
    const float gv = scale * d;
    dy[(size_t)m * n + k] = gv;
    dyb[(size_t)m * n + k] = f2bf(gv);
  }
  __shared__ float red[kEw];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kEw / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[m] = red[0] / (float)n;
}

// dz = dh * gelu'(z)
__global__ void k_gelu_bwd(const float* dh, const float* z, int rows, int cols, float* dz,
                           unsigned short* dzb) {
  const int m = (int)blockIdx.y;
  const size_t n = (size_t)rows * cols;
  for (size_t k = blockIdx.x * (size_t)kEw + threadIdx.x; k < n; k += (size_t)gridDim.x * kEw) {
    const size_t q = (size_t)m * n + k;
    const float v = dh[q] * gelu_grad(z[q]);
    dz[q] = v;
    dzb[q] = f2bf(v);
  }
}

// gradient of a bias: column sums of d (rows x cols) in row order
__global__ void k_colsum(const float* d, int rows, int cols, float* g, size_t bias_off,
                         size_t nparam) {
  const int m = (int)blockIdx.y;
  const int c = (int)(blockIdx.x * kEw + threadIdx.x);
  if (c >= cols) return;
  const float* dm = d + (size_t)m * rows * cols;
  float acc = 0.0f;
  for (int r = 0; r < rows; r++) acc += dm[(size_t)r * cols + c];
  g[(size_t)m * nparam + bias_off + c] = acc;
}

__global__ void k_to_bf16(const float* x, size_t n, unsigned short* y) {
  for (size_t k = blockIdx.x * (size_t)kEw + threadIdx.x; k < n; k += (size_t)gridDim.x * kEw)
    y[k] = f2bf(x[k]);
}

// gather rows idx[m][b] of the data set into the member's batch
__global__ void k_gather(const float* src, const int* idx, int rows, int cols, float* dst) {
  const int m = (int)blockIdx.y;
  const size_t n = (size_t)rows * cols;
  for (size_t k = blockIdx.x * (size_t)kEw + threadIdx.x; k < n; k += (size_t)gridDim.x * kEw) {
    const int r = (int)(k / cols), c = (int)(k % cols);
    dst[(size_t)m * n + k] = src[(size_t)idx[(size_t)m * rows + r] * cols + c];
  }
}

// Adam (Kingma & Ba 2015) on the FP32 master weights, then the BF16 copy.
__global__ void k_adam(float* p, const float* g, float* m1, float* m2, unsigned short* wb,
                       size_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (size_t k = blockIdx.x * (size_t)kEw + threadIdx.x; k < n; k += (size_t)gridDim.x * kEw) {
    const float gk = g[k];
    const float a = b1 * m1[k] + (1.0f - b1) * gk;
    const float b = b2 * m2[k] + (1.0f - b2) * gk * gk;
    m1[k] = a;
    m2[k] = b;
    const float v = p[k] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    p[k] = v;
    wb[k] = f2bf(v);
  }
}

int ew_blocks(size_t n) { return (int)std::min<size_t>(1024, (n + kEw - 1) / kEw); }

void gemm(bd_mlp* s, const unsigned short* A, long sAm, long sAi, long sAk,
          const unsigned short* B, long sBm, long sBk, long sBj, float* C, long sCm, long sCi,
          long sCj, int M, int N, int K) {
  GemmArgs g{A, sAm, sAi, sAk, B, sBm, sBk, sBj, C, sCm, sCi, sCj, M, N, K};
  dim3 grid((N + 31) / 32, (M + 31) / 32, s->members);
  k_gemm_bf16<<<grid, 64, 0, s->stream>>>(g);
}

int ensure_capacity(bd_mlp* s, int batch) {
  if (batch <= s->cap) return BD_OK;
  float** fb[] = {&s->x, &s->z1, &s->z2, &s->y, &s->dy, &s->dh, &s->dz, &s->dz1, &s->t};
  unsigned short** hb[] = {&s->xb, &s->h1b, &s->h2b, &s->dyb, &s->dz2b, &s->dz1b};
  for (float** q : fb)
    if (*q) (void)hipFree(*q), *q = nullptr;
  for (unsigned short** q : hb)
    if (*q) (void)hipFree(*q), *q = nullptr;
  const size_t M = (size_t)s->members, B = (size_t)batch;
  const size_t wmax = (size_t)std::max(std::max(s->n_in, s->hidden), s->n_out);
  MLP_TRY(hipMalloc((void**)&s->x, M * B * s->n_in * sizeof(float)));
  MLP_TRY(hipMalloc((void**)&s->t, M * B * s->n_out * sizeof(float)));
  MLP_TRY(hipMalloc((void**)&s->z1, M * B * s->hidden * sizeof(float)));
  MLP_TRY(hipMalloc((void**)&s->z2, M * B * s->hidden * sizeof(float)));
  MLP_TRY(hipMalloc((void**)&s->y, M * B * s->n_out * sizeof(float)));
  MLP_TRY(hipMalloc((void**)&s->dy, M * B * s->n_out * sizeof(float)));
  MLP_TRY(hipMalloc((void**)&s->dh, M * B * wmax * sizeof(float)));
  MLP_TRY(hipMalloc((void**)&s->dz, M * B * wmax * sizeof(float)));
  MLP_TRY(hipMalloc((void**)&s->dz1, M * B * s->hidden * sizeof(float)));
  MLP_TRY(hipMalloc((void**)&s->xb, M * B * s->n_in * sizeof(unsigned short)));
  MLP_TRY(hipMalloc((void**)&s->h1b, M * B * s->hidden * sizeof(unsigned short)));
  MLP_TRY(hipMalloc((void**)&s->h2b, M * B * s->hidden * sizeof(unsigned short)));
  MLP_TRY(hipMalloc((void**)&s->dyb, M * B * s->n_out * sizeof(unsigned short)));
  MLP_TRY(hipMalloc((void**)&s->dz2b, M * B * s->hidden * sizeof(unsigned short)));
  MLP_TRY(hipMalloc((void**)&s->dz1b, M * B * s->hidden * sizeof(unsigned short)));
  s->cap = batch;
  return BD_OK;
}

// Forward of every member on the member batches s->x[m] (batch rows).
void forward(bd_mlp* s, int batch) {
  const int M = s->members, H = s->hidden, I = s->n_in, O = s->n_out;
  const long P = (long)s->nparam;
  k_to_bf16<<<ew_blocks((size_t)M * batch * I), kEw, 0, s->stream>>>(
      s->x, (size_t)M * batch * I, s->xb);
  // layer 1: z1 = x W1 + b1, h1 = gelu(z1)
  gemm(s, s->xb, (long)batch * I, I, 1, s->wb + s->off_w1, P, H, 1, s->z1, (long)batch * H, H, 1,
       batch, H, I);
  k_bias_gelu<<<dim3(ew_blocks((size_t)batch * H), M), kEw, 0, s->stream>>>(
      s->z1, s->p, s->off_b1, s->nparam, batch, H, s->z1, s->h1b);
  gemm(s, s->h1b, (long)batch * H, H, 1, s->wb + s->off_w2, P, H, 1, s->z2, (long)batch * H, H, 1,
       batch, H, H);
  k_bias_gelu<<<dim3(ew_blocks((size_t)batch * H), M), kEw, 0, s->stream>>>(
      s->z2, s->p, s->off_b2, s->nparam, batch, H, s->z2, s->h2b);
  gemm(s, s->h2b, (long)batch * H, H, 1, s->wb + s->off_w3, P, O, 1, s->y, (long)batch * O, O, 1,
       batch, O, H);
  k_bias<<<dim3(ew_blocks((size_t)batch * O), M), kEw, 0, s->stream>>>(s->y, s->p, s->off_b3,
                                                                        s->nparam, batch, O);
}

// Backward of the mean-squared error against s->t (per member targets);
// writes the gradients of every member into s->g.  The FP32 gradients of the
// pre-activations stay in s->dz (layer 2) and s->dz1 (layer 1), their BF16
// copies in s->dz2b and s->dz1b (read back by bd_mlp_trace).
// The members' losses go to loss_out[0..members).
void backward(bd_mlp* s, int batch, float* loss_out) {
  const int M = s->members, H = s->hidden, I = s->n_in, O = s->n_out;
  const long P = (long)s->nparam;
  k_loss_grad<<<M, kEw, 0, s->stream>>>(s->y, s->t, 1, batch, O, s->dy, s->dyb, loss_out);
  // W3: h2^T dy ; b3: colsum(dy)
  gemm(s, s->h2b, (long)batch * H, 1, H, s->dyb, (long)batch * O, O, 1, s->g + s->off_w3, P, O, 1,
       H, O, batch);
  k_colsum<<<dim3((O + kEw - 1) / kEw, M), kEw, 0, s->stream>>>(s->dy, batch, O, s->g, s->off_b3,
                                                                s->nparam);
  // dh2 = dy W3^T
  gemm(s, s->dyb, (long)batch * O, O, 1, s->wb + s->off_w3, P, 1, O, s->dh, (long)batch * H, H, 1,
       batch, H, O);
  k_gelu_bwd<<<dim3(ew_blocks((size_t)batch * H), M), kEw, 0, s->stream>>>(s->dh, s->z2, batch,
                                                                           H, s->dz, s->dz2b);
  gemm(s, s->h1b, (long)batch * H, 1, H, s->dz2b, (long)batch * H, H, 1, s->g + s->off_w2, P, H,
       1, H, H, batch);
  k_colsum<<<dim3((H + kEw - 1) / kEw, M), kEw, 0, s->stream>>>(s->dz, batch, H, s->g, s->off_b2,
                                                                s->nparam);
  // dh1 = dz2 W2^T
  gemm(s, s->dz2b, (long)batch * H, H, 1, s->wb + s->off_w2, P, 1, H, s->dh, (long)batch * H, H,
       1, batch, H, H);
  k_gelu_bwd<<<dim3(ew_blocks((size_t)batch * H), M), kEw, 0, s->stream>>>(s->dh, s->z1, batch,
                                                                           H, s->dz1, s->dz1b);
  gemm(s, s->xb, (long)batch * I, 1, I, s->dz1b, (long)batch * H, H, 1, s->g + s->off_w1, P, H,
       1, I, H, batch);
  k_colsum<<<dim3((H + kEw - 1) / kEw, M), kEw, 0, s->stream>>>(s->dz1, batch, H, s->g,
                                                                s->off_b1, s->nparam);
}

void adam(bd_mlp* s, float lr) {
  s->adam_t++;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  // bias corrections in FP64: 1 - 0.999^t in FP32 would lose ~1e-5 to
  // cancellation in the first steps
  const float c1 = (float)(1.0 - std::pow(0.9, (double)s->adam_t));
  const float c2 = (float)(1.0 - std::pow(0.999, (double)s->adam_t));
  const size_t n = (size_t)s->members * s->nparam;
  k_adam<<<ew_blocks(n), kEw, 0, s->stream>>>(s->p, s->g, s->m1, s->m2, s->wb, n, lr, b1, b2,
                                               eps, c1, c2);
}

// splitmix64 stream for initialisation and shuffles (Steele et al. 2014)
struct Mix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

}  // namespace

extern "C" {

int bd_mlp_create(int32_t members, int32_t n_in, int32_t hidden, int32_t n_out, uint64_t seed,
                  bd_mlp** out) {
  if (!out || members < 1 || members > 64 || n_in < 1 || hidden < 1 || n_out < 1)
    return bdi::set_error(BD_E_ARG, "bad MLP shape");
  bd_mlp* s = new bd_mlp();
  s->members = members;
  s->n_in = n_in;
  s->hidden = hidden;
  s->n_out = n_out;
  s->off_w1 = 0;
  s->off_b1 = s->off_w1 + (size_t)n_in * hidden;
  s->off_w2 = s->off_b1 + hidden;
  s->off_b2 = s->off_w2 + (size_t)hidden * hidden;
  s->off_w3 = s->off_b2 + hidden;
  s->off_b3 = s->off_w3 + (size_t)hidden * n_out;
  s->nparam = s->off_b3 + n_out;
  const size_t n = (size_t)members * s->nparam;
  // He-style uniform initialisation, biases zero, one stream per member
  std::vector<float> h(n, 0.0f);
  for (int m = 0; m < members; m++) {
    Mix r{seed * 0x100000001b3ULL + (uint64_t)m};
    float* pm = &h[(size_t)m * s->nparam];
    auto fill = [&](size_t off, int fan_in, size_t count) {
      const double a = std::sqrt(6.0 / fan_in);
      for (size_t k = 0; k < count; k++) pm[off + k] = (float)((2.0 * r.uniform() - 1.0) * a);
    };
    fill(s->off_w1, n_in, (size_t)n_in * hidden);
    fill(s->off_w2, hidden, (size_t)hidden * hidden);
    fill(s->off_w3, hidden, (size_t)hidden * n_out);
  }
  hipError_t e = hipMalloc((void**)&s->p, n * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&s->g, n * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&s->m1, n * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&s->m2, n * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&s->wb, n * sizeof(unsigned short));
  if (e == hipSuccess) e = hipMalloc((void**)&s->loss, 64 * sizeof(float));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMemcpy(s->p, h.data(), n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(s->g, 0, n * sizeof(float));
  if (e == hipSuccess) e = hipMemset(s->m1, 0, n * sizeof(float));
  if (e == hipSuccess) e = hipMemset(s->m2, 0, n * sizeof(float));
  if (e == hipSuccess) {
    k_to_bf16<<<ew_blocks(n), kEw, 0, s->stream>>>(s->p, n, s->wb);
    e = hipStreamSynchronize(s->stream);
  }
  if (e != hipSuccess) {
    bd_mlp_destroy(s);
    return bdi::set_error(BD_E_HIP, std::string("bd_mlp_create: ") + hipGetErrorString(e));
  }
  *out = s;
  return BD_OK;
}

int bd_mlp_destroy(bd_mlp* s) {
  if (!s) return BD_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  void* ptrs[] = {s->p,  s->g,   s->m1,  s->m2,  s->wb,   s->loss, s->x,    s->t,  s->z1,
                  s->z2, s->y,   s->dy,  s->dh,  s->dz,   s->dz1,  s->xb,   s->h1b, s->h2b,
                  s->dyb, s->dz2b, s->dz1b, s->dx, s->dt, s->idx, s->lhist};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return BD_OK;
}

int bd_mlp_num_params(const bd_mlp* s, int64_t* out) {
  if (!s || !out) return bdi::set_error(BD_E_ARG, "null argument");
  *out = (int64_t)s->nparam;
  return BD_OK;
}

// params: members * nparam FP32 values in the packing documented above
int bd_mlp_get_params(bd_mlp* s, float* params) {
  if (!s || !params) return bdi::set_error(BD_E_ARG, "null argument");
  MLP_TRY(hipMemcpy(params, s->p, (size_t)s->members * s->nparam * sizeof(float),
                    hipMemcpyDeviceToHost));
  return BD_OK;
}

int bd_mlp_set_params(bd_mlp* s, const float* params) {
  if (!s || !params) return bdi::set_error(BD_E_ARG, "null argument");
  const size_t n = (size_t)s->members * s->nparam;
  MLP_TRY(hipMemcpy(s->p, params, n * sizeof(float), hipMemcpyHostToDevice));
  k_to_bf16<<<ew_blocks(n), kEw, 0, s->stream>>>(s->p, n, s->wb);
  MLP_TRY(hipStreamSynchronize(s->stream));
  return BD_OK;
}

// y[m][b][o] for batch rows x[b][i] (the same input for every member)
int bd_mlp_forward(bd_mlp* s, const float* x, int32_t batch, float* y) {
  if (!s || !x || !y || batch < 1) return bdi::set_error(BD_E_ARG, "bad arguments");
  int rc = ensure_capacity(s, batch);
  if (rc != BD_OK) return rc;
  const size_t xi = (size_t)batch * s->n_in;
  for (int m = 0; m < s->members; m++)
    MLP_TRY(hipMemcpyAsync(s->x + m * xi, x, xi * sizeof(float), hipMemcpyHostToDevice,
                           s->stream));
  forward(s, batch);
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipMemcpyAsync(y, s->y, (size_t)s->members * batch * s->n_out * sizeof(float),
                         hipMemcpyDeviceToHost, s->stream));
  MLP_TRY(hipStreamSynchronize(s->stream));
  return BD_OK;
}

// Gradients of the mean-squared error of every member on one batch (same
// inputs and targets for all members): grads[m * nparam + ...]; loss[m].
int bd_mlp_gradients(bd_mlp* s, const float* x, const float* t, int32_t batch, float* grads,
                     float* loss) {
  if (!s || !x || !t || batch < 1) return bdi::set_error(BD_E_ARG, "bad arguments");
  int rc = ensure_capacity(s, batch);
  if (rc != BD_OK) return rc;
  const size_t xi = (size_t)batch * s->n_in, to = (size_t)batch * s->n_out;
  for (int m = 0; m < s->members; m++) {
    MLP_TRY(hipMemcpyAsync(s->x + m * xi, x, xi * sizeof(float), hipMemcpyHostToDevice,
                           s->stream));
    MLP_TRY(hipMemcpyAsync(s->t + m * to, t, to * sizeof(float), hipMemcpyHostToDevice,
                           s->stream));
  }
  forward(s, batch);
  backward(s, batch, s->loss);
  MLP_TRY(hipGetLastError());
  if (grads)
    MLP_TRY(hipMemcpyAsync(grads, s->g, (size_t)s->members * s->nparam * sizeof(float),
                           hipMemcpyDeviceToHost, s->stream));
  if (loss)
    MLP_TRY(hipMemcpyAsync(loss, s->loss, (size_t)s->members * sizeof(float),
                           hipMemcpyDeviceToHost, s->stream));
  MLP_TRY(hipStreamSynchronize(s->stream));
  return BD_OK;
}

// The intermediates of member `member` from the last bd_mlp_gradients (or
// bd_mlp_train_step) call on `batch` rows, as floats, one array after the
// other: z1, h1, z2, h2 [batch][hidden]; y, dy, dyb [batch][n_out]; dz2,
// dz2b, dz1, dz1b [batch][hidden] (h*, dyb, dz*b: the BF16 operands of the
// matrix cores; z*, dy, dz*: FP32).
int bd_mlp_trace(bd_mlp* s, int32_t member, int32_t batch, float* out) {
  if (!s || !out || member < 0 || member >= s->members || batch < 1 || batch > s->cap)
    return bdi::set_error(BD_E_ARG, "bad arguments");
  const size_t bh = (size_t)batch * s->hidden, bo = (size_t)batch * s->n_out;
  const size_t mh = (size_t)member * bh, mo = (size_t)member * bo;
  std::vector<unsigned short> u(bh);
  auto f32 = [&](const float* src, size_t n, float* dst) -> int {
    MLP_TRY(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    return BD_OK;
  };
  auto b16 = [&](const unsigned short* src, size_t n, float* dst) -> int {
    MLP_TRY(hipMemcpyAsync(u.data(), src, n * sizeof(unsigned short), hipMemcpyDeviceToHost,
                           s->stream));
    MLP_TRY(hipStreamSynchronize(s->stream));
    for (size_t k = 0; k < n; k++) {
      const unsigned int w = (unsigned int)u[k] << 16;
      std::memcpy(&dst[k], &w, 4);
    }
    return BD_OK;
  };
  float* o = out;
  int rc = f32(s->z1 + mh, bh, o);
  if (rc == BD_OK) rc = b16(s->h1b + mh, bh, o + bh);
  if (rc == BD_OK) rc = f32(s->z2 + mh, bh, o + 2 * bh);
  if (rc == BD_OK) rc = b16(s->h2b + mh, bh, o + 3 * bh);
  o += 4 * bh;
  if (rc == BD_OK) rc = f32(s->y + mo, bo, o);
  if (rc == BD_OK) rc = f32(s->dy + mo, bo, o + bo);
  if (rc == BD_OK) rc = b16(s->dyb + mo, bo, o + 2 * bo);
  o += 3 * bo;
  if (rc == BD_OK) rc = f32(s->dz + mh, bh, o);
  if (rc == BD_OK) rc = b16(s->dz2b + mh, bh, o + bh);
  if (rc == BD_OK) rc = f32(s->dz1 + mh, bh, o + 2 * bh);
  if (rc == BD_OK) rc = b16(s->dz1b + mh, bh, o + 3 * bh);
  if (rc == BD_OK) MLP_TRY(hipStreamSynchronize(s->stream));
  return rc;
}

// Device-resident training set: x[n][n_in], t[n][n_out].
int bd_mlp_set_data(bd_mlp* s, const float* x, const float* t, int32_t n) {
  if (!s || !x || !t || n < 1) return bdi::set_error(BD_E_ARG, "bad arguments");
  if (s->dx) (void)hipFree(s->dx), s->dx = nullptr;
  if (s->dt) (void)hipFree(s->dt), s->dt = nullptr;
  MLP_TRY(hipMalloc((void**)&s->dx, (size_t)n * s->n_in * sizeof(float)));
  MLP_TRY(hipMalloc((void**)&s->dt, (size_t)n * s->n_out * sizeof(float)));
  MLP_TRY(hipMemcpy(s->dx, x, (size_t)n * s->n_in * sizeof(float), hipMemcpyHostToDevice));
  MLP_TRY(hipMemcpy(s->dt, t, (size_t)n * s->n_out * sizeof(float), hipMemcpyHostToDevice));
  s->n_data = n;
  return BD_OK;
}

// Adam training of all members on the device-resident set: `epochs` passes,
// each member with its own shuffle (seeded), mini-batches of `batch` rows.
// losses[epoch * members + m] = mean batch loss of the epoch.
int bd_mlp_train(bd_mlp* s, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                 float* losses) {
  if (!s || epochs < 1 || batch < 1) return bdi::set_error(BD_E_ARG, "bad arguments");
  if (s->n_data < batch) return bdi::set_error(BD_E_STATE, "set a data set of >= batch rows");
  int rc = ensure_capacity(s, batch);
  if (rc != BD_OK) return rc;
  const int M = s->members, nb = s->n_data / batch;
  if (s->idx) (void)hipFree(s->idx), s->idx = nullptr;
  if (s->lhist) (void)hipFree(s->lhist), s->lhist = nullptr;
  MLP_TRY(hipMalloc((void**)&s->idx, (size_t)M * s->n_data * sizeof(int)));
  MLP_TRY(hipMalloc((void**)&s->lhist, (size_t)M * nb * sizeof(float)));
  std::vector<int> perm((size_t)M * s->n_data);
  std::vector<float> lsum((size_t)M), lb((size_t)M * nb);
  Mix r{seed};
  for (int ep = 0; ep < epochs; ep++) {
    for (int m = 0; m < M; m++) {
      int* pm = &perm[(size_t)m * s->n_data];
      for (int k = 0; k < s->n_data; k++) pm[k] = k;
      for (int k = s->n_data - 1; k > 0; k--) {
        const int j = (int)(r.next() % (uint64_t)(k + 1));
        std::swap(pm[k], pm[j]);
      }
    }
    MLP_TRY(hipMemcpyAsync(s->idx, perm.data(), perm.size() * sizeof(int),
                           hipMemcpyHostToDevice, s->stream));
    std::fill(lsum.begin(), lsum.end(), 0.0f);
    for (int b = 0; b < nb; b++) {
      // member m's batch b: rows idx[m][b*batch ...]; the gather reads a
      // per-member row list with stride n_data
      for (int m = 0; m < M; m++) {
        const int* im = s->idx + (size_t)m * s->n_data + (size_t)b * batch;
        k_gather<<<dim3(ew_blocks((size_t)batch * s->n_in), 1), kEw, 0, s->stream>>>(
            s->dx, im, batch, s->n_in, s->x + (size_t)m * batch * s->n_in);
        k_gather<<<dim3(ew_blocks((size_t)batch * s->n_out), 1), kEw, 0, s->stream>>>(
            s->dt, im, batch, s->n_out, s->t + (size_t)m * batch * s->n_out);
      }
      forward(s, batch);
      backward(s, batch, s->lhist + (size_t)b * M);
      adam(s, lr);
    }
    // one read-back per epoch; it also keeps the host's shuffle of the next
    // epoch from overwriting perm before this epoch's copy has landed
    MLP_TRY(hipMemcpyAsync(lb.data(), s->lhist, lb.size() * sizeof(float),
                           hipMemcpyDeviceToHost, s->stream));
    MLP_TRY(hipStreamSynchronize(s->stream));
    for (int b = 0; b < nb; b++)
      for (int m = 0; m < M; m++) lsum[m] += lb[(size_t)b * M + m];
    if (losses)
      for (int m = 0; m < M; m++) losses[(size_t)ep * M + m] = lsum[m] / (float)nb;
  }
  MLP_TRY(hipGetLastError());
  return BD_OK;
}

// One Adam step of every member on a host batch (same rows for all members).
int bd_mlp_train_step(bd_mlp* s, const float* x, const float* t, int32_t batch, float lr,
                      float* loss) {
  int rc = bd_mlp_gradients(s, x, t, batch, nullptr, loss);
  if (rc != BD_OK) return rc;
  adam(s, lr);
  MLP_TRY(hipStreamSynchronize(s->stream));
  return BD_OK;
}

// Ensemble prediction: mean and (population) standard deviation over the
// members, n rows processed in chunks of up to 1024.
int bd_mlp_predict(bd_mlp* s, const float* x, int32_t n, float* mean, float* stdev) {
  if (!s || !x || !mean || !stdev || n < 1) return bdi::set_error(BD_E_ARG, "bad arguments");
  const int chunk = 1024;
  std::vector<float> y((size_t)s->members * chunk * s->n_out);
  for (int r0 = 0; r0 < n; r0 += chunk) {
    const int b = std::min(chunk, n - r0);
    int rc = bd_mlp_forward(s, x + (size_t)r0 * s->n_in, b, y.data());
    if (rc != BD_OK) return rc;
    for (int i = 0; i < b; i++)
      for (int o = 0; o < s->n_out; o++) {
        double a = 0.0, q = 0.0;
        for (int m = 0; m < s->members; m++) {
          const double v = y[((size_t)m * b + i) * s->n_out + o];
          a += v;
          q += v * v;
        }
        a /= s->members;
        q = q / s->members - a * a;
        mean[(size_t)(r0 + i) * s->n_out + o] = (float)a;
        stdev[(size_t)(r0 + i) * s->n_out + o] = (float)std::sqrt(q > 0.0 ? q : 0.0);
      }
  }
  return BD_OK;
}

}  // extern "C"
