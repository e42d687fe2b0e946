// The rank model: a small multilayer perceptron that predicts a box's
// skeleton count from its features, with hand-written forward, backward and
// Adam kernels (FP64).  Layout of the parameter vector (nparams doubles):
//
//   W1 (H x F, row-major), b1 (H), W2 (H x H), b2 (H), w3 (H), b3 (1)
//
// y = w3 . tanh(W2 tanh(W1 x + b1) + b2) + b3, loss = mean (y - t)^2 over a
// batch.  F = kMlpIn features per sample (row-major n x F), H = kMlpHidden.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "devutil.hpp"

namespace rskx {

constexpr int kMlpIn = 8;
constexpr int kMlpHidden = 32;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden +
                           kMlpHidden + 1;
constexpr int kMlpThreads = 256;
constexpr int kMlpGradRows = 32;   // samples per workgroup of the gradient kernel

int mlp_nparams() { return kMlpParams; }

namespace {
struct MlpView {
  const double *W1, *b1, *W2, *b2, *w3, *b3;
};
__device__ inline MlpView mlp_view(const double* p) {
  MlpView v;
  v.W1 = p;
  v.b1 = v.W1 + kMlpHidden * kMlpIn;
  v.W2 = v.b1 + kMlpHidden;
  v.b2 = v.W2 + kMlpHidden * kMlpHidden;
  v.w3 = v.b2 + kMlpHidden;
  v.b3 = v.w3 + kMlpHidden;
  return v;
}

// one sample's forward pass: activations a1, a2 (H each) and the output
__device__ inline double mlp_forward1(const MlpView& v, const double* x, double* a1, double* a2) {
  for (int h = 0; h < kMlpHidden; ++h) {
    double s = v.b1[h];
#pragma unroll
    for (int f = 0; f < kMlpIn; ++f) s += v.W1[h * kMlpIn + f] * x[f];
    a1[h] = tanh(s);
  }
  double y = v.b3[0];
  for (int h = 0; h < kMlpHidden; ++h) {
    double s = v.b2[h];
    for (int g = 0; g < kMlpHidden; ++g) s += v.W2[h * kMlpHidden + g] * a1[g];
    a2[h] = tanh(s);
    y += v.w3[h] * a2[h];
  }
  return y;
}
}  // namespace

// y[i] = model(X[i, :]); parameters staged in shared memory
__global__ void __launch_bounds__(kMlpThreads) k_mlp_forward(const double* __restrict__ params,
                                                             const double* __restrict__ X,
                                                             int64_t n, double* __restrict__ y) {
  __shared__ double p[kMlpParams];
  for (int e = threadIdx.x; e < kMlpParams; e += blockDim.x) p[e] = params[e];
  __syncthreads();
  const MlpView v = mlp_view(p);
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[kMlpIn], a1[kMlpHidden], a2[kMlpHidden];
#pragma unroll
  for (int f = 0; f < kMlpIn; ++f) x[f] = X[i * kMlpIn + f];
  y[i] = mlp_forward1(v, x, a1, a2);
}

// Gradient of sum_i (y_i - t_i)^2 / n with respect to the parameters:
// each workgroup back-propagates kMlpGradRows samples (one thread each,
// the samples' inputs, activations and deltas kept in shared memory), then
// sums the outer products dW1 = d1 x', dW2 = d2 a1', dw3 = d3 a2 and the
// biases over its samples into its row of `part` (nblocks x nparams); the
// loss contribution goes to lpart[block].  Summed over blocks in a fixed
// order by k_mlp_reduce, so the result does not depend on scheduling.
__global__ void __launch_bounds__(kMlpThreads) k_mlp_grad(const double* __restrict__ params,
                                                          const double* __restrict__ X,
                                                          const double* __restrict__ t, int64_t n,
                                                          double* __restrict__ part,
                                                          double* __restrict__ lpart) {
  constexpr int R = kMlpGradRows, H = kMlpHidden, F = kMlpIn;
  __shared__ double p[kMlpParams];
  __shared__ double sx[R][F], sa1[R][H], sa2[R][H], sd1[R][H], sd2[R][H], sd3[R];
  __shared__ double red[2 * kMlpThreads / 64];
  for (int e = threadIdx.x; e < kMlpParams; e += blockDim.x) p[e] = params[e];
  __syncthreads();
  const MlpView v = mlp_view(p);
  const int64_t i0 = blockIdx.x * (int64_t)R;
  const int r = threadIdx.x;
  double loss = 0;
  if (r < R) {
    const int64_t i = i0 + r;
    double x[F], a1[H], a2[H];
    if (i < n) {
#pragma unroll
      for (int f = 0; f < F; ++f) x[f] = X[i * F + f];
      const double y = mlp_forward1(v, x, a1, a2);
      const double e = y - t[i];
      loss = e * e;
      const double d3 = 2.0 * e / (double)n;
      sd3[r] = d3;
      for (int h = 0; h < H; ++h) sd2[r][h] = d3 * v.w3[h] * (1.0 - a2[h] * a2[h]);
      for (int g = 0; g < H; ++g) {
        double s = 0;
        for (int h = 0; h < H; ++h) s += v.W2[h * H + g] * sd2[r][h];
        sd1[r][g] = s * (1.0 - a1[g] * a1[g]);
      }
    } else {
#pragma unroll
      for (int f = 0; f < F; ++f) x[f] = 0;
      for (int h = 0; h < H; ++h) { a1[h] = 0; a2[h] = 0; sd1[r][h] = 0; sd2[r][h] = 0; }
      sd3[r] = 0;
    }
#pragma unroll
    for (int f = 0; f < F; ++f) sx[r][f] = x[f];
    for (int h = 0; h < H; ++h) { sa1[r][h] = a1[h]; sa2[r][h] = a2[h]; }
  }
  loss = block_sum(loss, red);
  __syncthreads();
  double* out = part + (int64_t)blockIdx.x * kMlpParams;
  for (int e = threadIdx.x; e < kMlpParams; e += blockDim.x) {
    double s = 0;
    if (e < H * F) {                                   // W1[h][f]
      const int h = e / F, f = e % F;
      for (int q = 0; q < R; ++q) s += sd1[q][h] * sx[q][f];
    } else if (e < H * F + H) {                        // b1
      const int h = e - H * F;
      for (int q = 0; q < R; ++q) s += sd1[q][h];
    } else if (e < H * F + H + H * H) {                // W2[h][g]
      const int o = e - H * F - H, h = o / H, g = o % H;
      for (int q = 0; q < R; ++q) s += sd2[q][h] * sa1[q][g];
    } else if (e < H * F + 2 * H + H * H) {            // b2
      const int h = e - H * F - H - H * H;
      for (int q = 0; q < R; ++q) s += sd2[q][h];
    } else if (e < H * F + 3 * H + H * H) {            // w3
      const int h = e - H * F - 2 * H - H * H;
      for (int q = 0; q < R; ++q) s += sd3[q] * sa2[q][h];
    } else {                                           // b3
      for (int q = 0; q < R; ++q) s += sd3[q];
    }
    out[e] = s;
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = loss;
}

// g[e] = sum over blocks of part[b][e] (fixed order); loss = sum lpart / n
__global__ void k_mlp_reduce(const double* __restrict__ part, const double* __restrict__ lpart,
                             int nblk, int64_t n, double* __restrict__ g, double* __restrict__ loss) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < kMlpParams) {
    double s = 0;
    for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * kMlpParams + e];
    g[e] = s;
  }
  if (e == 0) {
    double s = 0;
    for (int b = 0; b < nblk; ++b) s += lpart[b];
    *loss = s / (double)n;
  }
}

// one Adam step (Kingma and Ba, 2015) at step count `step` (>= 1)
__global__ void k_mlp_adam(double* __restrict__ params, const double* __restrict__ g,
                           double* __restrict__ m, double* __restrict__ v, double lr, double b1,
                           double b2, double eps, int step) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kMlpParams) return;
  const double ge = g[e];
  const double me = b1 * m[e] + (1.0 - b1) * ge;
  const double ve = b2 * v[e] + (1.0 - b2) * ge * ge;
  m[e] = me;
  v[e] = ve;
  const double mh = me / (1.0 - pow(b1, (double)step));
  const double vh = ve / (1.0 - pow(b2, (double)step));
  params[e] -= lr * mh / (sqrt(vh) + eps);
}

// ------------------------------------------------------------ launchers --
int mlp_grad_blocks(int64_t n) { return (int)((n + kMlpGradRows - 1) / kMlpGradRows); }

void launch_mlp_forward(const double* params, const double* X, int64_t n, double* y,
                        hipStream_t st) {
  if (n <= 0) return;
  const int grid = (int)((n + kMlpThreads - 1) / kMlpThreads);
  hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(kMlpThreads), 0, st, params, X, n, y);
}

void launch_mlp_grad(const double* params, const double* X, const double* t, int64_t n,
                     double* part, double* lpart, double* g, double* loss, hipStream_t st) {
  const int nblk = mlp_grad_blocks(n);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(kMlpThreads), 0, st, params, X, t, n, part,
                     lpart);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, part, lpart,
                     nblk, n, g, loss);
}

void launch_mlp_adam(double* params, const double* g, double* m, double* v, double lr, double b1,
                     double b2, double eps, int step, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, params, g, m, v,
                     lr, b1, b2, eps, step);
}

}  // namespace rskx

namespace rskx {

// Point moments of boxes for the rank model's features: box i owns the
// tree-order range [beg[i], beg[i] + cnt[i]); out[5 i ..] = count, sum x,
// sum y, sum z, sum |x|^2 (one workgroup per box).
__global__ void __launch_bounds__(256) k_box_moments(const int64_t* __restrict__ beg,
                                                     const int64_t* __restrict__ cnt,
                                                     const double* __restrict__ x0,
                                                     const double* __restrict__ x1,
                                                     const double* __restrict__ x2,
                                                     double* __restrict__ out) {
  __shared__ double red[8];
  const int64_t b0 = beg[blockIdx.x], c = cnt[blockIdx.x];
  double s[4] = {0, 0, 0, 0};
  for (int64_t q = threadIdx.x; q < c; q += blockDim.x) {
    const int64_t p = b0 + q;
    const double a = x0[p], bb = x1[p], cc = x2 ? x2[p] : 0.0;
    s[0] += a;
    s[1] += bb;
    s[2] += cc;
    s[3] += a * a + bb * bb + cc * cc;
  }
  for (int j = 0; j < 4; ++j) {
    const double t = block_sum(s[j], red);
    if (threadIdx.x == 0) out[5 * (int64_t)blockIdx.x + 1 + j] = t;
  }
  if (threadIdx.x == 0) out[5 * (int64_t)blockIdx.x] = (double)c;
}

void launch_box_moments(const int64_t* beg, const int64_t* cnt, int nbox, const double* x0,
                        const double* x1, const double* x2, double* out, hipStream_t st) {
  if (nbox > 0)
    hipLaunchKernelGGL(k_box_moments, dim3(nbox), dim3(256), 0, st, beg, cnt, x0, x1, x2, out);
}

}  // namespace rskx
