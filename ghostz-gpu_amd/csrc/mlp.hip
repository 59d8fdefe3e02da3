// mlp.hip -- see mlp.h.
//
// Every contraction runs through one bf16 MFMA GEMM kernel
// (v_mfma_f32_32x32x8bf16_1k, fp32 accumulation): a 256-thread block owns a
// 64 x 64 tile of C = A * B^T, each wave a 32 x 32 sub-tile; A and B are
// addressed through (row, k) strides so the transposed products of the
// backward pass reuse the same kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "cost_model.h"
#include "gz_common.h"
#include "mlp.h"

namespace gz {

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

constexpr size_t kOffEmbed = 0;
constexpr size_t kOffW1 = kOffEmbed + (size_t)kCostAlphabet * kCostEmbed;
constexpr size_t kOffB1 = kOffW1 + (size_t)kCostH1 * kCostIn;
constexpr size_t kOffW2 = kOffB1 + kCostH1;
constexpr size_t kOffB2 = kOffW2 + (size_t)kCostH2 * kCostH1;
constexpr size_t kOffW3 = kOffB2 + kCostH2;
constexpr size_t kOffB3 = kOffW3 + kCostH2;
constexpr size_t kFlat = kOffB3 + 1;

__device__ __forceinline__ uint16_t F2Bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float Bf2F(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float Gelu(float z) {
  return 0.5f * z * (1.f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float GeluGrad(float z) {
  const float cdf = 0.5f * (1.f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

enum { kEpiStore = 0, kEpiBiasGelu = 1, kEpiGeluGrad = 2 };

// C[m][n] = sum_k A(m, k) * B(n, k), A(m,k) = A[m*sam + k*sak], B likewise.
template <int EPI>
__global__ void __launch_bounds__(256)
    GemmKernel(int M, int N, int K, const uint16_t *A, long sam, long sak,
               const uint16_t *B, long sbn, long sbk, float *C, int ldc,
               const float *bias, uint16_t *Ch, const float *zsaved) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m0 = blockIdx.y * 64 + (wave >> 1) * 32;
  const int n0 = blockIdx.x * 64 + (wave & 1) * 32;
  const int am = m0 + (lane & 31), bn = n0 + (lane & 31);
  const int kq = (lane >> 5) * 4;
  float16_t acc;
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  for (int k0 = 0; k0 < K; k0 += 8) {
    short4_t a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + kq + j;
      a[j] = (am < M && k < K) ? (short)A[am * sam + k * sak] : (short)0;
      b[j] = (bn < N && k < K) ? (short)B[bn * sbn + k * sbk] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = m0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int col = n0 + (lane & 31);
    if (row < M && col < N) {
      const long idx = (long)row * ldc + col;
      const float v = acc[r];
      if (EPI == kEpiStore) {
        C[idx] = v;
      } else if (EPI == kEpiBiasGelu) {
        const float z = v + bias[col];
        C[idx] = z;
        Ch[idx] = F2Bf(Gelu(z));
      } else {
        const float dz = v * GeluGrad(zsaved[idx]);
        C[idx] = dz;
        Ch[idx] = F2Bf(dz);
      }
    }
  }
}

__global__ void EmbedKernel(int B, const uint8_t *codes, const uint16_t *embh, uint16_t *x) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)B * kCostIn) return;
  const long b = i / kCostIn;
  const int c = (int)(i % kCostIn);
  const int r = c / kCostEmbed, e = c % kCostEmbed;
  const int a = min((int)codes[b * kCostResidues + r], kCostAlphabet - 1);
  x[i] = embh[a * kCostEmbed + e];
}

__global__ void OutputKernel(int B, const uint16_t *h2, const float *w3, const float *b3,
                             float *y) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float s = 0.f;
  for (int k = 0; k < kCostH2; ++k) s += Bf2F(h2[b * kCostH2 + k]) * Bf2F(F2Bf(w3[k]));
  y[b] = s + b3[0];
}

// MSE loss over the first n rows; dY = 2 (y - t) / n, zero on padding.
__global__ void LossKernel(int B, int n, const float *y, const float *t, float *dy,
                           float *loss) {
  __shared__ float part[256];
  float acc = 0.f;
  for (int b = threadIdx.x; b < B; b += blockDim.x) {
    float g = 0.f;
    if (b < n) {
      const float d = y[b] - t[b];
      acc += d * d;
      g = 2.f * d / n;
    }
    dy[b] = g;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = part[0] / n;
}

// dW3 / dB3 and dZ2 = dY w3 * gelu'(Z2)
__global__ void OutputBackwardKernel(int B, const float *dy, const uint16_t *h2,
                                     const float *z2, const float *w3, float *gw3,
                                     float *gb3, float *dz2, uint16_t *dz2h) {
  const int k = threadIdx.x;  // 64 threads
  float s = 0.f, sb = 0.f;
  const float w = Bf2F(F2Bf(w3[k]));
  for (int b = 0; b < B; ++b) {
    s += dy[b] * Bf2F(h2[b * kCostH2 + k]);
    sb += dy[b];
    const float d = dy[b] * w * GeluGrad(z2[b * kCostH2 + k]);
    dz2[b * kCostH2 + k] = d;
    dz2h[b * kCostH2 + k] = F2Bf(d);
  }
  gw3[k] = s;
  if (k == 0) gb3[0] = sb;
}

// column sums of a [B][N] matrix into grad[N]
__global__ void ColumnSumKernel(int B, int N, const float *m, float *out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float s = 0.f;
  for (int b = 0; b < B; ++b) s += m[(long)b * N + c];
  out[c] = s;
}

// d_embed[a][e] = sum over (b, r) with code a of dX[b][r*8 + e]
__global__ void EmbedBackwardKernel(int B, const uint8_t *codes, const float *dx, float *gemb) {
  __shared__ float acc[kCostAlphabet * kCostEmbed];
  for (int i = threadIdx.x; i < kCostAlphabet * kCostEmbed; i += blockDim.x) acc[i] = 0.f;
  __syncthreads();
  const long total = (long)B * kCostIn;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const long b = i / kCostIn;
    const int c = (int)(i % kCostIn);
    const int a = min((int)codes[b * kCostResidues + c / kCostEmbed], kCostAlphabet - 1);
    atomicAdd(&acc[a * kCostEmbed + c % kCostEmbed], dx[i]);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kCostAlphabet * kCostEmbed; i += blockDim.x)
    atomicAdd(&gemb[i], acc[i]);
}

__global__ void AdamKernel(size_t n, float *w, const float *g, float *m, float *v, float lr,
                           float b1, float b2, float c1, float c2, float eps) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void ToBf16Kernel(size_t n, const float *src, uint16_t *dst) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = F2Bf(src[i]);
}

inline unsigned Blocks(size_t n, unsigned t = 256) { return (unsigned)((n + t - 1) / t); }

template <int EPI>
void Gemm(hipStream_t s, int M, int N, int K, const uint16_t *A, long sam, long sak,
          const uint16_t *B, long sbn, long sbk, float *C, int ldc, const float *bias = nullptr,
          uint16_t *Ch = nullptr, const float *z = nullptr) {
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  hipLaunchKernelGGL((GemmKernel<EPI>), grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbn,
                     sbk, C, ldc, bias, Ch, z);
}

}  // namespace

MlpDevice::MlpDevice() {
  (void)hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking);
  count_ = kFlat;
  (void)hipMalloc((void **)&master_, kFlat * 4);
  (void)hipMalloc((void **)&grad_, kFlat * 4);
  (void)hipMalloc((void **)&m_, kFlat * 4);
  (void)hipMalloc((void **)&v_, kFlat * 4);
  (void)hipMemset(m_, 0, kFlat * 4);
  (void)hipMemset(v_, 0, kFlat * 4);
  (void)hipMalloc((void **)&w1h_, (size_t)kCostH1 * kCostIn * 2);
  (void)hipMalloc((void **)&w2h_, (size_t)kCostH2 * kCostH1 * 2);
  (void)hipMalloc((void **)&embh_, (size_t)kCostAlphabet * kCostEmbed * 2);
}

MlpDevice::~MlpDevice() {
  void *ps[] = {master_, grad_, m_,   v_,    w1h_,  w2h_,  embh_, codes_, x_,
                z1_,     h1_,   z2_,  h2_,   y_,    t_,    loss_, dz2_,   dz2h_,
                dz1_,    dz1h_, dh1_, dx_};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void MlpDevice::SyncBf16() {
  hipLaunchKernelGGL(ToBf16Kernel, dim3(Blocks((size_t)kCostH1 * kCostIn)), dim3(256), 0,
                     stream_, (size_t)kCostH1 * kCostIn, master_ + kOffW1, w1h_);
  hipLaunchKernelGGL(ToBf16Kernel, dim3(Blocks((size_t)kCostH2 * kCostH1)), dim3(256), 0,
                     stream_, (size_t)kCostH2 * kCostH1, master_ + kOffW2, w2h_);
  hipLaunchKernelGGL(ToBf16Kernel, dim3(1), dim3(256), 0, stream_,
                     (size_t)kCostAlphabet * kCostEmbed, master_ + kOffEmbed, embh_);
}

void MlpDevice::SetWeights(const CostModelWeights &w) {
  std::vector<float> flat(kFlat);
  float *f = flat.data();
  auto put = [&](const std::vector<float> &v, size_t off) { memcpy(f + off, v.data(), v.size() * 4); };
  put(w.embed, kOffEmbed);
  put(w.w1, kOffW1);
  put(w.b1, kOffB1);
  put(w.w2, kOffW2);
  put(w.b2, kOffB2);
  put(w.w3, kOffW3);
  put(w.b3, kOffB3);
  (void)hipMemcpyAsync(master_, flat.data(), kFlat * 4, hipMemcpyHostToDevice, stream_);
  (void)hipMemsetAsync(m_, 0, kFlat * 4, stream_);
  (void)hipMemsetAsync(v_, 0, kFlat * 4, stream_);
  step_ = 0;
  SyncBf16();
  (void)hipStreamSynchronize(stream_);
}

void MlpDevice::GetWeights(CostModelWeights *w) {
  std::vector<float> flat(kFlat);
  (void)hipMemcpy(flat.data(), master_, kFlat * 4, hipMemcpyDeviceToHost);
  auto get = [&](std::vector<float> &v, size_t off, size_t n) {
    v.assign(flat.begin() + off, flat.begin() + off + n);
  };
  get(w->embed, kOffEmbed, (size_t)kCostAlphabet * kCostEmbed);
  get(w->w1, kOffW1, (size_t)kCostH1 * kCostIn);
  get(w->b1, kOffB1, kCostH1);
  get(w->w2, kOffW2, (size_t)kCostH2 * kCostH1);
  get(w->b2, kOffB2, kCostH2);
  get(w->w3, kOffW3, kCostH2);
  get(w->b3, kOffB3, 1);
}

hipError_t MlpDevice::Reserve(size_t batch) {
  if (batch <= cap_) return hipSuccess;
  void **ps[] = {(void **)&codes_, (void **)&x_,   (void **)&z1_,  (void **)&h1_,
                 (void **)&z2_,    (void **)&h2_,  (void **)&y_,   (void **)&t_,
                 (void **)&loss_,  (void **)&dz2_, (void **)&dz2h_, (void **)&dz1_,
                 (void **)&dz1h_,  (void **)&dh1_, (void **)&dx_};
  for (void **p : ps)
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  const size_t B = batch;
  hipError_t e = hipSuccess;
  auto al = [&](void **p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 64));
  };
  al((void **)&codes_, B * kCostResidues);
  al((void **)&x_, B * kCostIn * 2);
  al((void **)&z1_, B * kCostH1 * 4);
  al((void **)&h1_, B * kCostH1 * 2);
  al((void **)&z2_, B * kCostH2 * 4);
  al((void **)&h2_, B * kCostH2 * 2);
  al((void **)&y_, B * 4);
  al((void **)&t_, B * 4);
  al((void **)&loss_, 64);
  al((void **)&dz2_, B * kCostH2 * 4);
  al((void **)&dz2h_, B * kCostH2 * 2);
  al((void **)&dz1_, B * kCostH1 * 4);
  al((void **)&dz1h_, B * kCostH1 * 2);
  al((void **)&dh1_, B * kCostH1 * 4);
  al((void **)&dx_, B * kCostIn * 4);
  if (e == hipSuccess) cap_ = batch;
  return e;
}

hipError_t MlpDevice::ForwardDevice(size_t n, size_t B) {
  (void)n;
  hipStream_t s = stream_;
  hipLaunchKernelGGL(EmbedKernel, dim3(Blocks(B * kCostIn)), dim3(256), 0, s, (int)B, codes_,
                     embh_, x_);
  Gemm<kEpiBiasGelu>(s, (int)B, kCostH1, kCostIn, x_, kCostIn, 1, w1h_, kCostIn, 1, z1_, kCostH1,
                     master_ + kOffB1, h1_);
  Gemm<kEpiBiasGelu>(s, (int)B, kCostH2, kCostH1, h1_, kCostH1, 1, w2h_, kCostH1, 1, z2_,
                     kCostH2, master_ + kOffB2, h2_);
  hipLaunchKernelGGL(OutputKernel, dim3(Blocks(B)), dim3(256), 0, s, (int)B, h2_,
                     master_ + kOffW3, master_ + kOffB3, y_);
  return hipGetLastError();
}

hipError_t MlpDevice::Forward(const uint8_t *codes, size_t n, float *out) {
  const size_t chunk = 1 << 16;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  last_forward_ms = 0;
  for (size_t b0 = 0; b0 < n; b0 += chunk) {
    const size_t m = std::min(chunk, n - b0);
    const size_t B = (m + 63) / 64 * 64;
    hipError_t e = Reserve(B);
    if (e != hipSuccess) return e;
    (void)hipMemsetAsync(codes_, 20, B * kCostResidues, stream_);
    (void)hipMemcpyAsync(codes_, codes + b0 * kCostResidues, m * kCostResidues,
                         hipMemcpyHostToDevice, stream_);
    (void)hipEventRecord(e0, stream_);
    e = ForwardDevice(m, B);
    if (e != hipSuccess) return e;
    (void)hipEventRecord(e1, stream_);
    (void)hipMemcpyAsync(out + b0, y_, m * 4, hipMemcpyDeviceToHost, stream_);
    e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) return e;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    last_forward_ms += ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return hipSuccess;
}

hipError_t MlpDevice::TrainStep(const uint8_t *codes, const float *targets, size_t n,
                                float lr, float *loss, float *flat_grad) {
  const size_t B = (n + 63) / 64 * 64;
  hipError_t e = Reserve(B);
  if (e != hipSuccess) return e;
  hipStream_t s = stream_;
  (void)hipMemsetAsync(codes_, 20, B * kCostResidues, s);
  (void)hipMemcpyAsync(codes_, codes, n * kCostResidues, hipMemcpyHostToDevice, s);
  (void)hipMemsetAsync(t_, 0, B * 4, s);
  (void)hipMemcpyAsync(t_, targets, n * 4, hipMemcpyHostToDevice, s);
  e = ForwardDevice(n, B);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(LossKernel, dim3(1), dim3(256), 0, s, (int)B, (int)n, y_, t_, y_, loss_);
  // y_ now holds dY
  (void)hipMemsetAsync(grad_, 0, kFlat * 4, s);
  hipLaunchKernelGGL(OutputBackwardKernel, dim3(1), dim3(kCostH2), 0, s, (int)B, y_, h2_, z2_,
                     master_ + kOffW3, grad_ + kOffW3, grad_ + kOffB3, dz2_, dz2h_);
  // dW2 = dZ2^T H1
  Gemm<kEpiStore>(s, kCostH2, kCostH1, (int)B, dz2h_, 1, kCostH2, h1_, 1, kCostH1,
                  grad_ + kOffW2, kCostH1);
  hipLaunchKernelGGL(ColumnSumKernel, dim3(1), dim3(kCostH2), 0, s, (int)B, kCostH2, dz2_,
                     grad_ + kOffB2);
  // dZ1 = (dZ2 W2) * gelu'(Z1)
  Gemm<kEpiGeluGrad>(s, (int)B, kCostH1, kCostH2, dz2h_, kCostH2, 1, w2h_, 1, kCostH1, dz1_,
                     kCostH1, nullptr, dz1h_, z1_);
  // dW1 = dZ1^T X
  Gemm<kEpiStore>(s, kCostH1, kCostIn, (int)B, dz1h_, 1, kCostH1, x_, 1, kCostIn,
                  grad_ + kOffW1, kCostIn);
  hipLaunchKernelGGL(ColumnSumKernel, dim3(1), dim3(kCostH1), 0, s, (int)B, kCostH1, dz1_,
                     grad_ + kOffB1);
  // dX = dZ1 W1, then the embedding gradient
  Gemm<kEpiStore>(s, (int)B, kCostIn, kCostH1, dz1h_, kCostH1, 1, w1h_, 1, kCostIn, dx_,
                  kCostIn);
  hipLaunchKernelGGL(EmbedBackwardKernel, dim3(std::min<size_t>(Blocks(B * kCostIn), 1024)),
                     dim3(256), 0, s, (int)n, codes_, dx_, grad_ + kOffEmbed);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (lr != 0.f) {
    ++step_;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)step_), c2 = 1.f - std::pow(b2, (float)step_);
    hipLaunchKernelGGL(AdamKernel, dim3(Blocks(kFlat)), dim3(256), 0, s, kFlat, master_, grad_,
                       m_, v_, lr, b1, b2, c1, c2, 1e-8f);
    SyncBf16();
  }
  if (loss) (void)hipMemcpyAsync(loss, loss_, 4, hipMemcpyDeviceToHost, s);
  if (flat_grad) (void)hipMemcpyAsync(flat_grad, grad_, kFlat * 4, hipMemcpyDeviceToHost, s);
  return hipStreamSynchronize(s);
}

}  // namespace gz

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace gz;

struct gz_mlp_s {
  MlpDevice dev;
};

extern "C" {

size_t gz_mlp_weight_count(void) { return kFlat; }

gz_status gz_mlp_create(const float *flat, size_t count, gz_mlp_t *out) {
  GZ_REQUIRE(out && flat && count == kFlat, "gz_mlp_create: bad weights");
  CostModelWeights w;
  const float *f = flat;
  auto take = [&](std::vector<float> &v, size_t n) {
    v.assign(f, f + n);
    f += n;
  };
  take(w.embed, (size_t)kCostAlphabet * kCostEmbed);
  take(w.w1, (size_t)kCostH1 * kCostIn);
  take(w.b1, kCostH1);
  take(w.w2, (size_t)kCostH2 * kCostH1);
  take(w.b2, kCostH2);
  take(w.w3, kCostH2);
  take(w.b3, 1);
  gz_mlp_s *h = new gz_mlp_s();
  h->dev.SetWeights(w);
  *out = h;
  return GZ_OK;
}

gz_status gz_mlp_init_weights(uint64_t seed, float *flat) {
  CostModelWeights w;
  w.Init(seed);
  float *f = flat;
  for (auto *v : {&w.embed, &w.w1, &w.b1, &w.w2, &w.b2, &w.w3, &w.b3}) {
    memcpy(f, v->data(), v->size() * 4);
    f += v->size();
  }
  return GZ_OK;
}

gz_status gz_mlp_destroy(gz_mlp_t h) {
  delete h;
  return GZ_OK;
}

gz_status gz_mlp_predict(gz_mlp_t h, const uint8_t *codes, size_t n, float *out,
                         float *kernel_ms) {
  GZ_REQUIRE(h && (n == 0 || (codes && out)), "gz_mlp_predict: null argument");
  GZ_HIP_CHECK(h->dev.Forward(codes, n, out));
  if (kernel_ms) *kernel_ms = h->dev.last_forward_ms;
  return GZ_OK;
}

gz_status gz_mlp_train_step(gz_mlp_t h, const uint8_t *codes, const float *targets, size_t n,
                            float lr, float *loss, float *flat_grad) {
  GZ_REQUIRE(h && codes && targets && n > 0, "gz_mlp_train_step: bad argument");
  GZ_HIP_CHECK(h->dev.TrainStep(codes, targets, n, lr, loss, flat_grad));
  return GZ_OK;
}

gz_status gz_mlp_get_weights(gz_mlp_t h, float *flat) {
  GZ_REQUIRE(h && flat, "null argument");
  CostModelWeights w;
  h->dev.GetWeights(&w);
  float *f = flat;
  for (auto *v : {&w.embed, &w.w1, &w.b1, &w.w2, &w.b2, &w.w3, &w.b3}) {
    memcpy(f, v->data(), v->size() * 4);
    f += v->size();
  }
  return GZ_OK;
}

void gz_cost_features(const uint8_t *queries, const uint8_t *database, const uint32_t *qpos,
                      const uint32_t *dpos, size_t n, int reverse, uint8_t *out) {
  for (size_t i = 0; i < n; ++i)
    CostFeatures(queries, database, qpos[i], dpos[i], reverse != 0, out + i * kCostResidues);
}

}  // extern "C"
