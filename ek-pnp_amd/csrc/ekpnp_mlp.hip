// ekpnp_mlp.hip -- warm-start surrogate: BF16 MFMA multilayer perceptron with
// FP32 accumulation, FP32 master weights and Adam, all on the GPU.
//
// Layer l maps a_{l-1} [B][in_l] to z_l = a_{l-1} W_l^T + b_l [B][out_l];
// hidden layers store z_l (FP32, for GELU') and a_l = bf16(GELU(z_l)); the
// last layer is linear.  Backward (loss = mean (y - t)^2):
//   g_{L+1} = 2 (y - t) / (B n_out)
//   dW_l = g_l^T a_{l-1},  db_l = sum_b g_l,  g_{l-1} = (g_l W_l) * GELU'(z_{l-1})
// with g_l rounded to BF16 before it enters a product.  One GEMM kernel does
// all three product shapes: C[M][N] = sum_k A(m,k) B(n,k) with arbitrary
// strides, 64x64 tiles, four waves of 2x2 v_mfma_f32_16x16x32_bf16.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ekpnp_mlp.h"
#include "../../include/ekpnp_seed.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 32, KPAD = 8;  // LDS rows padded to 80 bytes
constexpr int GEMM_THREADS = 256;

// round-to-nearest-even FP32 -> BF16 bits
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}


__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float dgelu(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum Epi { EPI_HIDDEN = 0, EPI_OUT = 1, EPI_DGELU = 2, EPI_WGRAD = 3 };

struct GemmArgs {
  const uint16_t* A;
  long long sam, sak;
  const uint16_t* B;
  long long sbn, sbk;
  int M, N, K;
  const float* bias;   // [N]            (EPI_HIDDEN, EPI_OUT)
  float* Cf;           // FP32 out [M][ldc]
  uint16_t* Cb;        // BF16 out [M][ldc] (EPI_HIDDEN, EPI_DGELU)
  const float* zaux;   // pre-activation [M][ldc] (EPI_DGELU)
  long long ldc;
};

// Loads a 64 x 32 tile of X(r, k) = X[r*sr + k*sk] into LDS as [r][k].
__device__ inline void load_tile(uint16_t (*T)[TK + KPAD], const uint16_t* X, long long sr,
                                 long long sk, int r0, int k0, int R, int K) {
  const int t = threadIdx.x;
  if (sk == 1) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {  // consecutive lanes walk k
      const int idx = t * 8 + e, r = idx >> 5, kk = idx & 31;
      const int gr = r0 + r, gk = k0 + kk;
      T[r][kk] = (gr < R && gk < K) ? X[(long long)gr * sr + gk] : (uint16_t)0;
    }
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) {  // consecutive lanes walk r
      const int idx = t + e * GEMM_THREADS, kk = idx >> 6, r = idx & 63;
      const int gr = r0 + r, gk = k0 + kk;
      T[r][kk] = (gr < R && gk < K) ? X[(long long)gr * sr + (long long)gk * sk] : (uint16_t)0;
    }
  }
}

template <int EPI>
__global__ __launch_bounds__(GEMM_THREADS) void k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[TM][TK + KPAD];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[TN][TK + KPAD];
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = w >> 1, wn = w & 1;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < g.K; k0 += TK) {
    load_tile(As, g.A, g.sam, g.sak, m0, k0, g.M, g.K);
    load_tile(Bs, g.B, g.sbn, g.sbk, n0, k0, g.N, g.K);
    __syncthreads();
    // lane layout of v_mfma_f32_16x16x32_bf16: A row / B column = lane % 16,
    // k = 8 (lane / 16) + 0..7
    const int kq = (lane >> 4) * 8;
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const u16x8 va = *reinterpret_cast<const u16x8*>(&As[wm * 32 + i * 16 + (lane & 15)][kq]);
      af[i] = __builtin_bit_cast(bf16x8, va);
      const u16x8 vb = *reinterpret_cast<const u16x8*>(&Bs[wn * 32 + i * 16 + (lane & 15)][kq]);
      bfr[i] = __builtin_bit_cast(bf16x8, vb);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // C layout: column = lane % 16, rows 4 (lane / 16) + 0..3
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * 32 + i * 16 + (lane >> 4) * 4 + r;
        if (m >= g.M || n >= g.N) continue;
        const long long o = (long long)m * g.ldc + n;
        float v = acc[i][j][r];
        if (EPI == EPI_HIDDEN) {
          v += g.bias[n];
          g.Cf[o] = v;
          g.Cb[o] = f2bf(gelu(v));
        } else if (EPI == EPI_OUT) {
          g.Cf[o] = v + g.bias[n];
        } else if (EPI == EPI_DGELU) {
          v *= dgelu(g.zaux[o]);
          g.Cf[o] = v;
          g.Cb[o] = f2bf(v);
        } else {
          g.Cf[o] = v;
        }
      }
    }
  }
}

// x (FP32 [B][n_in]) -> BF16 [B][in_pad], zero padding
__global__ void k_pack_input(const float* __restrict__ x, uint16_t* __restrict__ xb, int B, int n_in,
                             int in_pad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * in_pad) return;
  const int b = i / in_pad, c = i - b * in_pad;
  xb[i] = (c < n_in) ? f2bf(x[(long long)b * n_in + c]) : (uint16_t)0;
}

// loss gradient g = 2 (y - t)/(B n), FP32 and BF16; per-block squared-error sums
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ y, const float* __restrict__ t,
                                              float* __restrict__ g32, uint16_t* __restrict__ gb,
                                              float* __restrict__ partial, int count, float scale) {
  __shared__ float red[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.f;
  if (i < count) {
    const float d = y[i] - t[i];
    const float gv = 2.0f * d * scale;
    g32[i] = gv;
    gb[i] = f2bf(gv);
    e2 = d * d;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) e2 += __shfl_down(e2, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e2;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// db[n] = sum_b g[b][n], fixed summation order
__global__ __launch_bounds__(256) void k_bias_grad(const float* __restrict__ g, float* __restrict__ db,
                                                   int B, int N) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int b = 0; b < B; ++b) s += g[(long long)b * N + n];
  db[n] = s;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ pb, long long n, float lr, float b1,
                       float b2, float eps, float wd, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * p[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  p[i] = pi;
  pb[i] = f2bf(pi);
}

__global__ void k_to_bf16(const float* __restrict__ p, uint16_t* __restrict__ pb, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pb[i] = f2bf(p[i]);
}

}  // namespace

struct ekpnp_mlp {
  ekpnp_mlp_config cfg{};
  int in_pad = 32, L = 0;
  std::vector<int> in_dim, out_dim;        // per layer 1..L+1 (index 0 unused)
  std::vector<long long> off_w, off_b;
  long long n_params = 0;
  float *p32 = nullptr, *m32 = nullptr, *v32 = nullptr, *g32 = nullptr;
  uint16_t* pbf = nullptr;
  uint16_t* xb = nullptr;
  std::vector<float*> z;       // [L+2], z[l] FP32 [B][out_l]
  std::vector<uint16_t*> a;    // [L+1], a[0] = xb, a[l] BF16 [B][out_l]
  float* t = nullptr;          // targets
  float* xin = nullptr;        // FP32 input staging
  float *gf = nullptr, *gf2 = nullptr;       // FP32 gradient rows
  uint16_t *gb = nullptr, *gb2 = nullptr;    // BF16 gradient rows
  float* partial = nullptr;
  int max_partials = 0;
  hipStream_t stream = nullptr;
  float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f, wd = 0.f;
  long long adam_t = 0;
  std::string err;
};

namespace {

int mfail(ekpnp_mlp* m, int code, const std::string& msg) {
  if (m) m->err = msg;
  return code;
}

#define MHIP(m, call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return mfail((m), -2, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

void release(ekpnp_mlp* m) {
  void* ptrs[] = {m->p32, m->m32, m->v32, m->g32, m->pbf, m->xb, m->t, m->xin, m->gf, m->gf2, m->gb, m->gb2,
                  m->partial};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (float* p : m->z)
    if (p) (void)hipFree(p);
  for (size_t i = 1; i < m->a.size(); ++i)
    if (m->a[i]) (void)hipFree(m->a[i]);
  if (m->stream) (void)hipStreamDestroy(m->stream);
}

hipError_t gemm(ekpnp_mlp* m, int epi, const GemmArgs& g) {
  const dim3 grid((g.N + TN - 1) / TN, (g.M + TM - 1) / TM), block(GEMM_THREADS);
  switch (epi) {
    case EPI_HIDDEN: hipLaunchKernelGGL(k_gemm<EPI_HIDDEN>, grid, block, 0, m->stream, g); break;
    case EPI_OUT: hipLaunchKernelGGL(k_gemm<EPI_OUT>, grid, block, 0, m->stream, g); break;
    case EPI_DGELU: hipLaunchKernelGGL(k_gemm<EPI_DGELU>, grid, block, 0, m->stream, g); break;
    default: hipLaunchKernelGGL(k_gemm<EPI_WGRAD>, grid, block, 0, m->stream, g); break;
  }
  return hipGetLastError();
}

hipError_t forward(ekpnp_mlp* m, const float* x_host, int B) {
  hipError_t e = hipMemcpyAsync(m->xin, x_host, (size_t)B * m->cfg.n_in * sizeof(float),
                                hipMemcpyHostToDevice, m->stream);
  if (e != hipSuccess) return e;
  const int n = B * m->in_pad;
  hipLaunchKernelGGL(k_pack_input, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->xin, m->xb, B,
                     m->cfg.n_in, m->in_pad);
  for (int l = 1; l <= m->L + 1; ++l) {
    GemmArgs g{};
    g.A = m->a[l - 1];
    g.sam = m->in_dim[l];
    g.sak = 1;
    g.B = m->pbf + m->off_w[l];
    g.sbn = m->in_dim[l];
    g.sbk = 1;
    g.M = B;
    g.N = m->out_dim[l];
    g.K = m->in_dim[l];
    g.bias = m->p32 + m->off_b[l];
    g.Cf = m->z[l];
    g.Cb = (l <= m->L) ? m->a[l] : nullptr;
    g.ldc = m->out_dim[l];
    e = gemm(m, l <= m->L ? EPI_HIDDEN : EPI_OUT, g);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// backward after forward: fills g32 with the gradient of the loss; loss sum
// into *loss (host)
int backward(ekpnp_mlp* m, const float* t_host, int B, float* loss) {
  const int L = m->L, no = m->cfg.n_out;
  MHIP(m, hipMemcpyAsync(m->t, t_host, (size_t)B * no * sizeof(float), hipMemcpyHostToDevice, m->stream));
  const int count = B * no, nblk = (count + 255) / 256;
  hipLaunchKernelGGL(k_loss, dim3(nblk), dim3(256), 0, m->stream, m->z[L + 1], m->t, m->gf, m->gb, m->partial,
                     count, 1.0f / (float)count);
  MHIP(m, hipGetLastError());
  float* gcur = m->gf;
  uint16_t* gbcur = m->gb;
  float* gnext = m->gf2;
  uint16_t* gbnext = m->gb2;
  for (int l = L + 1; l >= 1; --l) {
    const int in = m->in_dim[l], out = m->out_dim[l];
    // dW_l[o][i] = sum_b g[b][o] a_{l-1}[b][i]
    GemmArgs gw{};
    gw.A = gbcur;
    gw.sam = 1;
    gw.sak = out;
    gw.B = m->a[l - 1];
    gw.sbn = 1;
    gw.sbk = in;
    gw.M = out;
    gw.N = in;
    gw.K = B;
    gw.Cf = m->g32 + m->off_w[l];
    gw.ldc = in;
    MHIP(m, gemm(m, EPI_WGRAD, gw));
    hipLaunchKernelGGL(k_bias_grad, dim3((out + 255) / 256), dim3(256), 0, m->stream, gcur,
                       m->g32 + m->off_b[l], B, out);
    MHIP(m, hipGetLastError());
    if (l > 1) {
      // g_{l-1}[b][i] = sum_o g[b][o] W_l[o][i] * GELU'(z_{l-1}[b][i])
      GemmArgs gd{};
      gd.A = gbcur;
      gd.sam = out;
      gd.sak = 1;
      gd.B = m->pbf + m->off_w[l];
      gd.sbn = 1;
      gd.sbk = in;
      gd.M = B;
      gd.N = in;
      gd.K = out;
      gd.Cf = gnext;
      gd.Cb = gbnext;
      gd.zaux = m->z[l - 1];
      gd.ldc = in;
      MHIP(m, gemm(m, EPI_DGELU, gd));
      std::swap(gcur, gnext);
      std::swap(gbcur, gbnext);
    }
  }
  std::vector<float> part(nblk);
  MHIP(m, hipMemcpyAsync(part.data(), m->partial, nblk * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MHIP(m, hipStreamSynchronize(m->stream));
  double s = 0.0;
  for (float v : part) s += v;
  if (loss) *loss = (float)(s / (double)count);
  return 0;
}

}  // namespace

extern "C" {

int ekpnp_mlp_create(const ekpnp_mlp_config* cfg, ekpnp_mlp** out) {
  if (!cfg || !out) return -1;
  *out = nullptr;
  if (cfg->n_in < 1 || cfg->n_in > 32 || cfg->n_hidden < 32 || cfg->n_hidden % 32 || cfg->n_layers < 1 ||
      cfg->n_out < 32 || cfg->n_out % 32 || cfg->max_batch < 1)
    return -1;
  if (hipSetDevice(cfg->device) != hipSuccess) return -2;
  ekpnp_mlp* m = new ekpnp_mlp();
  m->cfg = *cfg;
  m->L = cfg->n_layers;
  const int L = m->L, H = cfg->n_hidden, B = cfg->max_batch;
  m->in_dim.assign(L + 2, 0);
  m->out_dim.assign(L + 2, 0);
  m->off_w.assign(L + 2, 0);
  m->off_b.assign(L + 2, 0);
  long long off = 0;
  for (int l = 1; l <= L + 1; ++l) {
    m->in_dim[l] = (l == 1) ? m->in_pad : H;
    m->out_dim[l] = (l <= L) ? H : cfg->n_out;
    m->off_w[l] = off;
    off += (long long)m->in_dim[l] * m->out_dim[l];
    m->off_b[l] = off;
    off += m->out_dim[l];
  }
  m->n_params = off;
  // He initialisation from the shared splitmix64 stream (Box-Muller)
  std::vector<float> p(off, 0.f);
  uint64_t st = ekpnp_stream_key(cfg->seed, 0x6d6c70ull, 0);
  for (int l = 1; l <= L + 1; ++l) {
    const int fan_in = (l == 1) ? cfg->n_in : H;
    const double sd = std::sqrt(2.0 / fan_in);
    for (int o = 0; o < m->out_dim[l]; ++o)
      for (int i = 0; i < m->in_dim[l]; ++i) {
        double u1 = ekpnp_uniform01(&st), u2 = ekpnp_uniform01(&st);
        double nrm = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * M_PI * u2);
        if (l == 1 && i >= cfg->n_in) nrm = 0.0;  // padding columns
        p[m->off_w[l] + (long long)o * m->in_dim[l] + i] = (float)(sd * nrm);
      }
  }
  const int maxw = std::max(H, cfg->n_out);
  m->max_partials = (B * cfg->n_out + 255) / 256;
  auto A = [&](void** ptr, size_t bytes) { return hipMalloc(ptr, std::max<size_t>(bytes, 16)); };
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = A((void**)&m->p32, off * 4);
  if (e == hipSuccess) e = A((void**)&m->m32, off * 4);
  if (e == hipSuccess) e = A((void**)&m->v32, off * 4);
  if (e == hipSuccess) e = A((void**)&m->g32, off * 4);
  if (e == hipSuccess) e = A((void**)&m->pbf, off * 2);
  if (e == hipSuccess) e = A((void**)&m->xb, (size_t)B * m->in_pad * 2);
  if (e == hipSuccess) e = A((void**)&m->xin, (size_t)B * cfg->n_in * 4);
  if (e == hipSuccess) e = A((void**)&m->t, (size_t)B * cfg->n_out * 4);
  if (e == hipSuccess) e = A((void**)&m->gf, (size_t)B * maxw * 4);
  if (e == hipSuccess) e = A((void**)&m->gf2, (size_t)B * maxw * 4);
  if (e == hipSuccess) e = A((void**)&m->gb, (size_t)B * maxw * 2);
  if (e == hipSuccess) e = A((void**)&m->gb2, (size_t)B * maxw * 2);
  if (e == hipSuccess) e = A((void**)&m->partial, (size_t)m->max_partials * 4);
  m->z.assign(L + 2, nullptr);
  m->a.assign(L + 1, nullptr);
  m->a[0] = m->xb;
  for (int l = 1; l <= L + 1 && e == hipSuccess; ++l) {
    e = A((void**)&m->z[l], (size_t)B * m->out_dim[l] * 4);
    if (e == hipSuccess && l <= L) e = A((void**)&m->a[l], (size_t)B * m->out_dim[l] * 2);
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMemcpy(m->p32, p.data(), off * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(m->m32, 0, off * 4);
  if (e == hipSuccess) e = hipMemset(m->v32, 0, off * 4);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((off + 255) / 256)), dim3(256), 0, m->stream, m->p32, m->pbf, off);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) {
    release(m);
    delete m;
    return e == hipErrorOutOfMemory ? -3 : -2;
  }
  *out = m;
  return 0;
}

void ekpnp_mlp_destroy(ekpnp_mlp* m) {
  if (!m) return;
  (void)hipSetDevice(m->cfg.device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  release(m);
  delete m;
}

const char* ekpnp_mlp_last_error(const ekpnp_mlp* m) { return m ? m->err.c_str() : "null surrogate"; }

long long ekpnp_mlp_num_params(const ekpnp_mlp* m) { return m ? m->n_params : 0; }

int ekpnp_mlp_get_params(ekpnp_mlp* m, float* out) {
  if (!m || !out) return -1;
  MHIP(m, hipSetDevice(m->cfg.device));
  MHIP(m, hipStreamSynchronize(m->stream));
  MHIP(m, hipMemcpy(out, m->p32, m->n_params * 4, hipMemcpyDeviceToHost));
  return 0;
}

int ekpnp_mlp_set_params(ekpnp_mlp* m, const float* in) {
  if (!m || !in) return -1;
  MHIP(m, hipSetDevice(m->cfg.device));
  MHIP(m, hipStreamSynchronize(m->stream));
  MHIP(m, hipMemcpy(m->p32, in, m->n_params * 4, hipMemcpyHostToDevice));
  MHIP(m, hipMemset(m->m32, 0, m->n_params * 4));
  MHIP(m, hipMemset(m->v32, 0, m->n_params * 4));
  m->adam_t = 0;
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((m->n_params + 255) / 256)), dim3(256), 0, m->stream, m->p32,
                     m->pbf, m->n_params);
  MHIP(m, hipGetLastError());
  MHIP(m, hipStreamSynchronize(m->stream));
  return 0;
}

int ekpnp_mlp_forward(ekpnp_mlp* m, const float* x, int batch, float* y) {
  if (!m || !x || !y) return -1;
  if (batch < 1 || batch > m->cfg.max_batch) return mfail(m, -1, "batch out of range");
  MHIP(m, hipSetDevice(m->cfg.device));
  MHIP(m, forward(m, x, batch));
  MHIP(m, hipMemcpyAsync(y, m->z[m->L + 1], (size_t)batch * m->cfg.n_out * 4, hipMemcpyDeviceToHost, m->stream));
  MHIP(m, hipStreamSynchronize(m->stream));
  return 0;
}

int ekpnp_mlp_loss_grad(ekpnp_mlp* m, const float* x, const float* t, int batch, float* loss, float* grad) {
  if (!m || !x || !t) return -1;
  if (batch < 1 || batch > m->cfg.max_batch) return mfail(m, -1, "batch out of range");
  MHIP(m, hipSetDevice(m->cfg.device));
  MHIP(m, forward(m, x, batch));
  int rc = backward(m, t, batch, loss);
  if (rc) return rc;
  if (grad) MHIP(m, hipMemcpy(grad, m->g32, m->n_params * 4, hipMemcpyDeviceToHost));
  return 0;
}

int ekpnp_mlp_set_adam(ekpnp_mlp* m, float beta1, float beta2, float eps, float weight_decay) {
  if (!m) return -1;
  m->b1 = beta1;
  m->b2 = beta2;
  m->eps = eps;
  m->wd = weight_decay;
  return 0;
}

int ekpnp_mlp_train_step(ekpnp_mlp* m, const float* x, const float* t, int batch, float lr, float* loss) {
  if (!m || !x || !t) return -1;
  if (batch < 1 || batch > m->cfg.max_batch) return mfail(m, -1, "batch out of range");
  MHIP(m, hipSetDevice(m->cfg.device));
  MHIP(m, forward(m, x, batch));
  int rc = backward(m, t, batch, loss);
  if (rc) return rc;
  m->adam_t++;
  const float c1 = 1.0f / (1.0f - std::pow(m->b1, (float)m->adam_t));
  const float c2 = 1.0f / (1.0f - std::pow(m->b2, (float)m->adam_t));
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->n_params + 255) / 256)), dim3(256), 0, m->stream, m->p32, m->g32,
                     m->m32, m->v32, m->pbf, m->n_params, lr, m->b1, m->b2, m->eps, m->wd, c1, c2);
  MHIP(m, hipGetLastError());
  return 0;
}

}  // extern "C"
