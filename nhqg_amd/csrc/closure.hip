// Learned vertical-tail closure: a per-column MLP acting on the top Chebyshev
// coefficients of q (no counterpart in the reference; its fallback and judge
// is the plain solver at higher Nz).
//
// Network: x (n_in) -> tanh(W1 x + b1) (width) -> tanh(W2 h1 + b2) (width)
// -> W3 h2 + b3 (n_out).  Parameters are one flat FP32 array in the order
// W1 (width x n_in, row-major), b1, W2 (width x width), b2, W3 (n_out x
// width), b3.  Training runs on the device with FP32 forward / backward
// kernels (sums in a fixed order, no atomics) and Adam; the step uses
// BF16-rounded weights with FP32 accumulation.
//
// A network created with NHQG_MLP_ODD keeps every bias at zero.  tanh is
// odd, so the whole map is then odd: zero columns get no tendency, and the
// k / -k columns at ky = 0 (q(-k) = conj q(k): equal real parts, opposite
// imaginary parts) get Hermitian tendencies.  Oddness is also what
// statistical translation invariance asks of a per-column predictor: a
// shift by half a wavelength maps a column to its negative.
#include <mutex>
#include <vector>
#include <string>
#include <cstring>
#include <cmath>
#include <algorithm>

#include "common.h"
#include "../../include/nhqg.h"

struct nhqg_mlp {
  int n_in = 0, width = 0, n_out = 0, max_batch = 0, flags = 0;
  size_t n_par = 0;
  float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;   // params, grads, Adam moments
  float *x = nullptr, *t = nullptr;                                  // batch input / target
  float *a1 = nullptr, *a2 = nullptr, *y = nullptr;                  // activations
  float *d1 = nullptr, *d2 = nullptr, *dy = nullptr;                 // deltas
  float* loss = nullptr;
  float* hist = nullptr;     // per-step loss of the last train call
  int hist_cap = 0;
  float* data_x = nullptr;   // resident training set
  float* data_t = nullptr;
  int n_data = 0;
  long adam_t = 0;
  hipStream_t st = nullptr;
  std::string err;
};

namespace {

size_t off_b1(const nhqg_mlp* m) { return (size_t)m->width * m->n_in; }
size_t off_w2(const nhqg_mlp* m) { return off_b1(m) + m->width; }
size_t off_b2(const nhqg_mlp* m) { return off_w2(m) + (size_t)m->width * m->width; }
size_t off_w3(const nhqg_mlp* m) { return off_b2(m) + m->width; }
size_t off_b3(const nhqg_mlp* m) { return off_w3(m) + (size_t)m->n_out * m->width; }

// out[b][o] = act(bias[o] + sum_i W[o][i] in[b][i]) (act = tanh or identity)
__global__ void __launch_bounds__(256)
mlp_dense_fwd(int batch, int n_in, int n_out, const float* __restrict__ W,
              const float* __restrict__ bias, const float* __restrict__ in, float* __restrict__ out,
              int act) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= batch * n_out) return;
  const int b = idx / n_out, o = idx % n_out;
  float s = bias[o];
  const float* wr = W + (size_t)o * n_in;
  const float* xr = in + (size_t)b * n_in;
  for (int i = 0; i < n_in; ++i) s = fmaf(wr[i], xr[i], s);
  out[idx] = act ? tanhf(s) : s;
}

// dy = 2 (y - t) / (batch * n_out); loss partial per block (fixed tree)
__global__ void __launch_bounds__(256)
mlp_loss(int n, const float* __restrict__ y, const float* __restrict__ t, float scale,
         float* __restrict__ dy, float* __restrict__ part) {
  __shared__ float red[256];
  const int idx = blockIdx.x * 256 + threadIdx.x;
  float e = 0.0f;
  if (idx < n) {
    const float d = y[idx] - t[idx];
    dy[idx] = 2.0f * d * scale;
    e = d * d;
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void mlp_loss_sum(int nb, float scale, float* part, float* hist) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float s = 0.0f;
    for (int i = 0; i < nb; ++i) s += part[i];
    part[nb] = s * scale;
    if (hist) *hist = s * scale;
  }
}

// zero the three bias ranges of a parameter-shaped array
__global__ void __launch_bounds__(256)
mlp_zero_bias(float* __restrict__ p, size_t o1, size_t o2, size_t o3, int width, int n_out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < width) {
    p[o1 + i] = 0.0f;
    p[o2 + i] = 0.0f;
  }
  if (i < n_out) p[o3 + i] = 0.0f;
}

// gradient of one layer from the delta at its output (pre-activation delta
// dpre[b][o]): gW[o][i] = sum_b dpre[b][o] in[b][i], gb[o] = sum_b dpre[b][o]
__global__ void __launch_bounds__(256)
mlp_dense_grad(int batch, int n_in, int n_out, const float* __restrict__ dpre,
               const float* __restrict__ in, float* __restrict__ gW, float* __restrict__ gb) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx < n_out * n_in) {
    const int o = idx / n_in, i = idx % n_in;
    float s = 0.0f;
    for (int b = 0; b < batch; ++b) s = fmaf(dpre[(size_t)b * n_out + o], in[(size_t)b * n_in + i], s);
    gW[idx] = s;
  } else if (idx < n_out * n_in + n_out) {
    const int o = idx - n_out * n_in;
    float s = 0.0f;
    for (int b = 0; b < batch; ++b) s += dpre[(size_t)b * n_out + o];
    gb[o] = s;
  }
}

// delta at a tanh layer's pre-activation from the next layer:
// dpre_in[b][i] = (1 - a[b][i]^2) * sum_o W[o][i] dpre_out[b][o]
__global__ void __launch_bounds__(256)
mlp_dense_back(int batch, int n_in, int n_out, const float* __restrict__ W,
               const float* __restrict__ dpre_out, const float* __restrict__ a,
               float* __restrict__ dpre_in) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= batch * n_in) return;
  const int b = idx / n_in, i = idx % n_in;
  float s = 0.0f;
  for (int o = 0; o < n_out; ++o) s = fmaf(W[(size_t)o * n_in + i], dpre_out[(size_t)b * n_out + o], s);
  const float av = a[idx];
  dpre_in[idx] = (1.0f - av * av) * s;
}

__global__ void __launch_bounds__(256)
mlp_adam(size_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
         float* __restrict__ m2, float lr, float b1, float b2, float c1, float c2, float eps) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  w[i] -= lr * (a * c1) / (sqrtf(v * c2) + eps);
}

__global__ void __launch_bounds__(256)
mlp_gather(int batch, int n_in, int n_out, int start, int n_data, const float* __restrict__ dx,
           const float* __restrict__ dt, float* __restrict__ x, float* __restrict__ t) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx < batch * n_in) {
    const int b = idx / n_in, i = idx % n_in;
    x[idx] = dx[(size_t)((start + b) % n_data) * n_in + i];
  }
  if (idx < batch * n_out) {
    const int b = idx / n_out, o = idx % n_out;
    t[idx] = dt[(size_t)((start + b) % n_data) * n_out + o];
  }
}

unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

void zero_bias(nhqg_mlp* m, float* p) {
  hipLaunchKernelGGL(mlp_zero_bias, dim3(nblk((size_t)std::max(m->width, m->n_out))), dim3(256), 0,
                     m->st, p, off_b1(m), off_b2(m), off_b3(m), m->width, m->n_out);
}

#define MK(expr)                                                              \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      m->err = std::string(#expr) + ": " + hipGetErrorString(e_);             \
      return NHQG_ERR_HIP;                                                    \
    }                                                                         \
  } while (0)

// The three layers on stream st: x -> a1 -> a2 -> y (all device memory).
int forward_into(nhqg_mlp* m, int batch, const float* x, float* a1, float* a2, float* y,
                 hipStream_t st) {
  const float* w = m->w;
  hipLaunchKernelGGL(mlp_dense_fwd, dim3(nblk((size_t)batch * m->width)), dim3(256), 0, st,
                     batch, m->n_in, m->width, w, w + off_b1(m), x, a1, 1);
  hipLaunchKernelGGL(mlp_dense_fwd, dim3(nblk((size_t)batch * m->width)), dim3(256), 0, st,
                     batch, m->width, m->width, w + off_w2(m), w + off_b2(m), a1, a2, 1);
  hipLaunchKernelGGL(mlp_dense_fwd, dim3(nblk((size_t)batch * m->n_out)), dim3(256), 0, st,
                     batch, m->width, m->n_out, w + off_w3(m), w + off_b3(m), a2, y, 0);
  MK(hipGetLastError());
  return NHQG_OK;
}

int forward(nhqg_mlp* m, int batch) {
  return forward_into(m, batch, m->x, m->a1, m->a2, m->y, m->st);
}

int backward(nhqg_mlp* m, int batch, float* hist = nullptr) {
  const int n = batch * m->n_out;
  const unsigned nb = nblk(n);
  const float sc = 1.0f / (float)n;
  hipLaunchKernelGGL(mlp_loss, dim3(nb), dim3(256), 0, m->st, n, m->y, m->t, sc, m->dy, m->loss);
  hipLaunchKernelGGL(mlp_loss_sum, dim3(1), dim3(64), 0, m->st, (int)nb, sc, m->loss, hist);
  float* g = m->g;
  const float* w = m->w;
  const int W = m->width;
  hipLaunchKernelGGL(mlp_dense_grad, dim3(nblk((size_t)m->n_out * W + m->n_out)), dim3(256), 0,
                     m->st, batch, W, m->n_out, m->dy, m->a2, g + off_w3(m), g + off_b3(m));
  hipLaunchKernelGGL(mlp_dense_back, dim3(nblk((size_t)batch * W)), dim3(256), 0, m->st, batch, W,
                     m->n_out, w + off_w3(m), m->dy, m->a2, m->d2);
  hipLaunchKernelGGL(mlp_dense_grad, dim3(nblk((size_t)W * W + W)), dim3(256), 0, m->st, batch, W,
                     W, m->d2, m->a1, g + off_w2(m), g + off_b2(m));
  hipLaunchKernelGGL(mlp_dense_back, dim3(nblk((size_t)batch * W)), dim3(256), 0, m->st, batch, W,
                     W, w + off_w2(m), m->d2, m->a1, m->d1);
  hipLaunchKernelGGL(mlp_dense_grad, dim3(nblk((size_t)W * m->n_in + W)), dim3(256), 0, m->st,
                     batch, m->n_in, W, m->d1, m->x, g, g + off_b1(m));
  MK(hipGetLastError());
  return NHQG_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Closure in the step: for every mode m and part (Re, Im), x_j = in_scale *
// q[n0 + j][m] (j < 16), E_q[n0 + j][m] += out_scale * MLP(x)_j with the odd
// (bias-free) network, on the matrix cores: the three layers are GEMMs over
// a batch of 16 columns per wavefront on v_mfma_f32_16x16x32_bf16 (BF16
// weights, FP32 accumulation).  Lane layout (checked on the device): A row =
// lane % 16, B column = lane % 16, k = 8 * (lane / 16) + e for the 8
// elements e of a lane, D row = 4 * (lane / 16) + r.
//
// A layer's D tiles feed the next layer's B fragments without any lane
// exchange: for k-step ks a lane's 8 elements are hidden units
//   u(ks, g, e) = 32 ks + 4 g + e (e < 4),  32 ks + 16 + 4 g + (e - 4) (e >= 4)
// of its own column (g = lane / 16), which are the values the lane holds in
// D tiles 2 ks and 2 ks + 1.  The weight fragments are permuted to that k
// order on the host (closure_export_bf16).  Activations enter the MFMAs as
// a BF16 pair hi + lo (about 16 bits), close to the FP32 activations the
// network was trained with.
// ---------------------------------------------------------------------------
#define CL_NIN 16
#define CL_W 128
// fragment buffer (bf16): W1 [8 tiles][64 lanes][8], W2 [8][4 k-steps][64][8],
// W3 [4 k-steps][64][8]
#define CL_F1 0
#define CL_F2 (8 * 64 * 8)
#define CL_F3 (CL_F2 + 8 * 4 * 64 * 8)
#define CL_FN (CL_F3 + 4 * 64 * 8)

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

NHQG_DEV void split_bf16(const float (&v)[8], v8bf& hi, v8bf& lo) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const __bf16 h = (__bf16)v[e];
    hi[e] = h;
    lo[e] = (__bf16)(v[e] - (float)h);
  }
}

__global__ void __launch_bounds__(256)
closure_mfma_kernel(int nm2, int ld, int n0, const v8bf* __restrict__ frag,
                    const double* __restrict__ q, double* __restrict__ Eq, float in_scale,
                    float out_scale) {
  __shared__ v8bf F[CL_FN / 8];
  for (int i = threadIdx.x; i < CL_FN / 8; i += 256) F[i] = frag[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4;
  const int col = blockIdx.x * 64 + wave * 16 + (lane & 15);
  const bool valid = col < nm2;
  const int cc = valid ? col : nm2 - 1;
  // layer 1 input: x_j, j = 8 g + e (g < 2), zero padding for g >= 2
  float xv[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int j = min(8 * g + e, CL_NIN - 1);
    const float x = in_scale * (float)q[(size_t)(n0 + j) * ld + cc];
    xv[e] = g < 2 ? x : 0.0f;
  }
  v8bf bh, bl;
  split_bf16(xv, bh, bl);
  v4f h1[8];
#pragma unroll
  for (int mt = 0; mt < 8; ++mt) {
    v4f a = {0.f, 0.f, 0.f, 0.f};
    const v8bf w = F[(CL_F1 / 8) + mt * 64 + lane];
    a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, bh, a, 0, 0, 0);
    a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, bl, a, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) h1[mt][r] = tanhf(a[r]);
  }
  // layer 2: B fragments of the four k-steps from the layer-1 tiles
  v8bf b2h[4], b2l[4];
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    float v[8];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      v[r] = h1[2 * ks][r];
      v[4 + r] = h1[2 * ks + 1][r];
    }
    split_bf16(v, b2h[ks], b2l[ks]);
  }
  v4f h2[8];
#pragma unroll
  for (int mt = 0; mt < 8; ++mt) {
    v4f a = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const v8bf w = F[(CL_F2 / 8) + (mt * 4 + ks) * 64 + lane];
      a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, b2h[ks], a, 0, 0, 0);
      a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, b2l[ks], a, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) h2[mt][r] = tanhf(a[r]);
  }
  // layer 3: one output tile, outputs 4 g + r of this lane's column
  v4f y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    float v[8];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      v[r] = h2[2 * ks][r];
      v[4 + r] = h2[2 * ks + 1][r];
    }
    v8bf ph, pl;
    split_bf16(v, ph, pl);
    const v8bf w = F[(CL_F3 / 8) + ks * 64 + lane];
    y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, ph, y, 0, 0, 0);
    y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, pl, y, 0, 0, 0);
  }
  if (valid) {
#pragma unroll
    for (int r = 0; r < 4; ++r) Eq[(size_t)(n0 + 4 * g + r) * ld + col] += (double)(out_scale * y[r]);
  }
}

hipError_t launch_closure(int nm2, int ld, int n0, const void* frag, const double* q, double* Eq,
                          float in_scale, float out_scale, hipStream_t s) {
  if (nm2 < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(closure_mfma_kernel, dim3((unsigned)((nm2 + 63) / 64)), dim3(256), 0, s, nm2,
                     ld, n0, (const v8bf*)frag, q, Eq, in_scale, out_scale);
  return hipGetLastError();
}

namespace {
// float -> bf16 bits, round to nearest even (finite inputs)
uint16_t to_bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
}  // namespace

// BF16 weight fragments of the step's closure in the kernel's LDS order
// (shapes must be 16 -> 128 -> 128 -> 16, odd network)
int closure_export_bf16(const nhqg_mlp* m, void** dst, hipStream_t s, std::string& err) {
  if (m->n_in != CL_NIN || m->width != CL_W || m->n_out != CL_NIN) {
    err = "the step closure needs a 16 -> 128 -> 128 -> 16 network";
    return NHQG_ERR_ARG;
  }
  if (!(m->flags & NHQG_MLP_ODD)) {
    err = "the step closure needs an odd network (created with NHQG_MLP_ODD)";
    return NHQG_ERR_ARG;
  }
  std::vector<float> w(m->n_par);
  if (hipStreamSynchronize(m->st) != hipSuccess ||
      hipMemcpy(w.data(), m->w, m->n_par * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
    err = "reading the closure weights failed";
    return NHQG_ERR_HIP;
  }
  const float* W1 = w.data();
  const float* W2 = W1 + CL_W * CL_NIN + CL_W;
  const float* W3 = W2 + CL_W * CL_W + CL_W;
  auto u = [](int ks, int g, int e) { return e < 4 ? 32 * ks + 4 * g + e : 32 * ks + 16 + 4 * g + e - 4; };
  std::vector<uint16_t> f(CL_FN, 0);
  for (int mt = 0; mt < 8; ++mt)
    for (int lane = 0; lane < 64; ++lane)
      for (int e = 0; e < 8; ++e) {
        const int j = 8 * (lane / 16) + e;
        f[CL_F1 + (mt * 64 + lane) * 8 + e] =
            j < CL_NIN ? to_bf16_bits(W1[(16 * mt + lane % 16) * CL_NIN + j]) : 0;
        for (int ks = 0; ks < 4; ++ks)
          f[CL_F2 + ((mt * 4 + ks) * 64 + lane) * 8 + e] =
              to_bf16_bits(W2[(16 * mt + lane % 16) * CL_W + u(ks, lane / 16, e)]);
      }
  for (int ks = 0; ks < 4; ++ks)
    for (int lane = 0; lane < 64; ++lane)
      for (int e = 0; e < 8; ++e)
        f[CL_F3 + (ks * 64 + lane) * 8 + e] = to_bf16_bits(W3[(lane % 16) * CL_W + u(ks, lane / 16, e)]);
  if (!*dst && hipMalloc(dst, CL_FN * sizeof(uint16_t)) != hipSuccess) {
    err = "hipMalloc failed for the closure weights";
    return NHQG_ERR_HIP;
  }
  if (hipMemcpyAsync(*dst, f.data(), CL_FN * sizeof(uint16_t), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    err = "uploading the closure weights failed";
    return NHQG_ERR_HIP;
  }
  return NHQG_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int nhqg_mlp_create(int32_t n_in, int32_t width, int32_t n_out, int32_t max_batch,
                    int32_t flags, nhqg_mlp** out) {
  if (!out || n_in < 1 || width < 1 || n_out < 1 || max_batch < 1 || (flags & ~NHQG_MLP_ODD))
    return NHQG_ERR_ARG;
  *out = nullptr;
  nhqg_mlp* m = new nhqg_mlp();
  m->flags = flags;
  m->n_in = n_in;
  m->width = width;
  m->n_out = n_out;
  m->max_batch = max_batch;
  m->n_par = (size_t)width * n_in + width + (size_t)width * width + width +
             (size_t)n_out * width + n_out;
  const size_t B = max_batch;
  bool ok = true;
  float** bufs[] = {&m->w, &m->g, &m->m1, &m->m2};
  for (float** b : bufs)
    ok = ok && hipMalloc(b, m->n_par * sizeof(float)) == hipSuccess &&
         hipMemset(*b, 0, m->n_par * sizeof(float)) == hipSuccess;
  size_t sz[] = {B * n_in, B * n_out, B * width, B * width, B * n_out, B * width, B * width,
                 B * n_out, (B * n_out + 255) / 256 + 1};
  float** act[] = {&m->x, &m->t, &m->a1, &m->a2, &m->y, &m->d1, &m->d2, &m->dy, &m->loss};
  for (int i = 0; i < 9 && ok; ++i) ok = hipMalloc(act[i], sz[i] * sizeof(float)) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) == hipSuccess;
  if (!ok) {
    nhqg_mlp_destroy(m);
    return NHQG_ERR_HIP;
  }
  *out = m;
  return NHQG_OK;
}

int nhqg_mlp_destroy(nhqg_mlp* m) {
  if (!m) return NHQG_OK;
  if (m->st) (void)hipStreamSynchronize(m->st);
  float* ptrs[] = {m->w, m->g, m->m1, m->m2, m->x, m->t, m->a1, m->a2, m->y, m->d1, m->d2,
                   m->dy, m->loss, m->hist, m->data_x, m->data_t};
  for (float* p : ptrs)
    if (p) (void)hipFree(p);
  if (m->st) (void)hipStreamDestroy(m->st);
  delete m;
  return NHQG_OK;
}

int nhqg_mlp_n_params(const nhqg_mlp* m, int64_t* n) {
  if (!m || !n) return NHQG_ERR_ARG;
  *n = (int64_t)m->n_par;
  return NHQG_OK;
}

int nhqg_mlp_set_params(nhqg_mlp* m, const float* w) {
  if (!m || !w) return NHQG_ERR_ARG;
  MK(hipStreamSynchronize(m->st));
  MK(hipMemcpy(m->w, w, m->n_par * sizeof(float), hipMemcpyHostToDevice));
  MK(hipMemset(m->m1, 0, m->n_par * sizeof(float)));
  MK(hipMemset(m->m2, 0, m->n_par * sizeof(float)));
  if (m->flags & NHQG_MLP_ODD) {
    zero_bias(m, m->w);
    MK(hipGetLastError());
    MK(hipStreamSynchronize(m->st));
  }
  m->adam_t = 0;
  return NHQG_OK;
}

int nhqg_mlp_get_params(nhqg_mlp* m, float* w) {
  if (!m || !w) return NHQG_ERR_ARG;
  MK(hipStreamSynchronize(m->st));
  MK(hipMemcpy(w, m->w, m->n_par * sizeof(float), hipMemcpyDeviceToHost));
  return NHQG_OK;
}

int nhqg_mlp_forward(nhqg_mlp* m, const float* x, int32_t batch, float* y) {
  if (!m) return NHQG_ERR_ARG;
  if (!x || !y || batch < 1 || batch > m->max_batch) {
    m->err = "nhqg_mlp_forward: null pointer, or batch outside [1, " +
             std::to_string(m->max_batch) + "]";
    return NHQG_ERR_ARG;
  }
  MK(hipMemcpyAsync(m->x, x, (size_t)batch * m->n_in * sizeof(float), hipMemcpyHostToDevice,
                    m->st));
  int r = forward(m, batch);
  if (r) return r;
  MK(hipMemcpyAsync(y, m->y, (size_t)batch * m->n_out * sizeof(float), hipMemcpyDeviceToHost,
                    m->st));
  MK(hipStreamSynchronize(m->st));
  return NHQG_OK;
}

int nhqg_mlp_gradient(nhqg_mlp* m, const float* x, const float* t, int32_t batch, float* grad,
                      float* loss) {
  if (!m) return NHQG_ERR_ARG;
  if (!x || !t || !grad || !loss || batch < 1 || batch > m->max_batch) {
    m->err = "nhqg_mlp_gradient: null pointer, or batch outside [1, " +
             std::to_string(m->max_batch) + "]";
    return NHQG_ERR_ARG;
  }
  MK(hipMemcpyAsync(m->x, x, (size_t)batch * m->n_in * sizeof(float), hipMemcpyHostToDevice,
                    m->st));
  MK(hipMemcpyAsync(m->t, t, (size_t)batch * m->n_out * sizeof(float), hipMemcpyHostToDevice,
                    m->st));
  int r = forward(m, batch);
  if (r) return r;
  r = backward(m, batch);
  if (r) return r;
  const unsigned nb = nblk((size_t)batch * m->n_out);
  MK(hipMemcpyAsync(grad, m->g, m->n_par * sizeof(float), hipMemcpyDeviceToHost, m->st));
  MK(hipMemcpyAsync(loss, m->loss + nb, sizeof(float), hipMemcpyDeviceToHost, m->st));
  MK(hipStreamSynchronize(m->st));
  return NHQG_OK;
}

int nhqg_mlp_set_data(nhqg_mlp* m, const float* x, const float* t, int32_t n) {
  if (!m || !x || !t || n < 1) return NHQG_ERR_ARG;
  if (m->data_x) (void)hipFree(m->data_x);
  if (m->data_t) (void)hipFree(m->data_t);
  m->data_x = m->data_t = nullptr;
  MK(hipMalloc(&m->data_x, (size_t)n * m->n_in * sizeof(float)));
  MK(hipMalloc(&m->data_t, (size_t)n * m->n_out * sizeof(float)));
  MK(hipMemcpy(m->data_x, x, (size_t)n * m->n_in * sizeof(float), hipMemcpyHostToDevice));
  MK(hipMemcpy(m->data_t, t, (size_t)n * m->n_out * sizeof(float), hipMemcpyHostToDevice));
  m->n_data = n;
  return NHQG_OK;
}

// n_steps of Adam over consecutive minibatches of the resident set (taken
// cyclically from `start`); loss[i] = loss of step i before its update.
int nhqg_mlp_train(nhqg_mlp* m, int32_t n_steps, int32_t batch, int32_t start, float lr,
                   float* loss) {
  if (!m || n_steps < 1 || batch < 1 || batch > m->max_batch || !m->data_x) return NHQG_ERR_ARG;
  if (n_steps > m->hist_cap) {
    MK(hipStreamSynchronize(m->st));
    if (m->hist) (void)hipFree(m->hist);
    m->hist = nullptr;
    m->hist_cap = 0;
    MK(hipMalloc(&m->hist, (size_t)n_steps * sizeof(float)));
    m->hist_cap = n_steps;
  }
  const bool odd = (m->flags & NHQG_MLP_ODD) != 0;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int s = 0; s < n_steps; ++s) {
    const int st = (int)(((long)start + (long)s * batch) % m->n_data);
    hipLaunchKernelGGL(mlp_gather, dim3(nblk((size_t)batch * std::max(m->n_in, m->n_out))),
                       dim3(256), 0, m->st, batch, m->n_in, m->n_out, st, m->n_data, m->data_x,
                       m->data_t, m->x, m->t);
    int r = forward(m, batch);
    if (r) return r;
    r = backward(m, batch, m->hist + s);
    if (r) return r;
    if (odd) zero_bias(m, m->g);
    m->adam_t += 1;
    const float c1 = 1.0f / (1.0f - powf(b1, (float)m->adam_t));
    const float c2 = 1.0f / (1.0f - powf(b2, (float)m->adam_t));
    hipLaunchKernelGGL(mlp_adam, dim3(nblk(m->n_par)), dim3(256), 0, m->st, m->n_par, m->w, m->g,
                       m->m1, m->m2, lr, b1, b2, c1, c2, eps);
  }
  MK(hipGetLastError());
  if (loss)
    MK(hipMemcpyAsync(loss, m->hist, (size_t)n_steps * sizeof(float), hipMemcpyDeviceToHost, m->st));
  MK(hipStreamSynchronize(m->st));
  return NHQG_OK;
}

const char* nhqg_mlp_last_error(const nhqg_mlp* m) { return m ? m->err.c_str() : ""; }

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free forward pass.  The two hidden activations live in one
// file-scope device array (4 MiB: NHQG_MLP_NOALLOC_MAX_BATCH rows of up to
// NHQG_MLP_NOALLOC_MAX_WIDTH floats, twice), held by every GPU the library
// runs on until the process ends.  Host code never passes the array's name
// to a kernel: its device address, taken once per device with that device
// current by hipGetSymbolAddress, is kept in g_scratch_ptr[device].
// ---------------------------------------------------------------------------

static __device__ __align__(256) float
    scratch[2 * NHQG_MLP_NOALLOC_MAX_BATCH * NHQG_MLP_NOALLOC_MAX_WIDTH];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(float** out, std::string& err) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    err = "nhqg_mlp_forward_noalloc: no usable current device";
    return NHQG_ERR_HIP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
      return NHQG_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return NHQG_OK;
}
}  // namespace

int nhqg_mlp_forward_noalloc(nhqg_mlp* m, const float* x, int32_t batch, float* y,
                             void* stream) {
  if (!m || !x || !y) return NHQG_ERR_ARG;
  if (batch < 1 || batch > NHQG_MLP_NOALLOC_MAX_BATCH) {
    m->err = "nhqg_mlp_forward_noalloc: batch must be in [1, " +
             std::to_string(NHQG_MLP_NOALLOC_MAX_BATCH) + "]";
    return NHQG_ERR_ARG;
  }
  if (m->width > NHQG_MLP_NOALLOC_MAX_WIDTH) {
    m->err = "nhqg_mlp_forward_noalloc: width must be at most " +
             std::to_string(NHQG_MLP_NOALLOC_MAX_WIDTH);
    return NHQG_ERR_ARG;
  }
  float* s = nullptr;
  const int r = scratch_for_current_device(&s, m->err);
  if (r) return r;
  // every element of both activations is written before it is read: nothing
  // in the scratch array needs clearing between calls
  float* a1 = s;
  float* a2 = s + (size_t)NHQG_MLP_NOALLOC_MAX_BATCH * NHQG_MLP_NOALLOC_MAX_WIDTH;
  return forward_into(m, batch, x, a1, a2, y, static_cast<hipStream_t>(stream));
}
