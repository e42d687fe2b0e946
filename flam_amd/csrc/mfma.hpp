// 16 x 16 x 4 matrix-core products shared by the factorization kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace rskx {

typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef double mf_d4 __attribute__((ext_vector_type(4)));

// 16 x 16 x 4 MFMA of each precision.  A (16 x 4): lane l holds A(l & 15,
// l >> 4); B (4 x 16): lane l holds B(l >> 4, l & 15); D (16 x 16): lane l
// holds column l & 15, its element v in row row(l, v) (the f32 and f64
// instructions order the rows differently).
template <typename S> struct Mf;
template <> struct Mf<float> {
  typedef mf_f4 V;
  __device__ static V mac(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <> struct Mf<double> {
  typedef mf_d4 V;
  __device__ static V mac(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

// workgroup barrier for LDS traffic only: LDS writes are complete before it,
// while global stores stay in flight (a full __syncthreads would wait for
// every outstanding store of the workgroup)
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// LDS written by some lanes of a wave and read by others of the same wave:
// a wave's LDS operations complete in order, so only the compiler must be
// kept from moving accesses across this point
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kMfPad = 4;                     // row padding of the LDS blocks (elements)
constexpr int kMfLdp = 16 + kMfPad;           // leading dimension of a panel row

// acc +/-= A B' over `inner` columns (a multiple of 16) of two 16-row
// row-major blocks (leading dimensions multiples of 4, rows 16-byte
// aligned for f32 and 32-byte aligned for f64).  Per 16 columns lane l
// reads the four consecutive entries 4 (l >> 4) + [0, 4) of row l & 15 of
// each block, and MFMA c multiplies their c-th entries: A and B pair the
// same inner index, and the four products cover all 16.
template <typename S, bool NEG>
__device__ inline typename Mf<S>::V mac_rows(typename Mf<S>::V acc, const S* A, int lda,
                                             const S* B, int ldb, int inner, int lane) {
  typedef typename Mf<S>::V V;
  const S* a = A + (lane & 15) * lda + 4 * (lane >> 4);
  const S* b = B + (lane & 15) * ldb + 4 * (lane >> 4);
  for (int c0 = 0; c0 < inner; c0 += 16) {
    const V av = *reinterpret_cast<const V*>(a + c0);
    const V bv = *reinterpret_cast<const V*>(b + c0);
#pragma unroll
    for (int c = 0; c < 4; ++c) acc = Mf<S>::mac(NEG ? -av[c] : av[c], bv[c], acc);
  }
  return acc;
}

}  // namespace rskx
