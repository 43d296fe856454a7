// Kernels of the learned warm-start network: a 3-level 3-D U-Net with BF16
// activations and weights, FP32 accumulation on the matrix cores (MFMA), FP32 master
// weights and Adam.
//
// Tensors are channels-last over the volume grid (Y, X, Z), z fastest, the layout of
// the reconstruction volume: element (v, c) of a tensor with channel stride cs is at
// v * cs + c, v = (y * X + x) * Z + z.  A channel slice of a wider tensor (the skip
// half of a concatenation) is addressed by offsetting the base pointer.
//
// 3x3x3 convolution ("same", zero padding) as an implicit GEMM:
//   out[v][co] = sum_{tap, ci} in[v + off(tap)][ci] * W[tap][co][ci] + bias[co]
// M = voxels, N = output channels, K = 27 * Cin flattened as k = tap * Cin + ci.
// One wave computes 32 voxels x Cout with v_mfma_f32_16x16x32_bf16.  Operand layout of
// that instruction (lane l, j < 8, r < 4): A[l % 16][8 (l / 16) + j],
// B[8 (l / 16) + j][l % 16], D[4 (l / 16) + r][l % 16].
// The backward-data pass is the same convolution of the output gradient with the
// flipped, channel-transposed weights Wb[tap][ci][co] = W[26 - tap][co][ci].
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "unet_internal.hpp"

namespace ctr {
namespace unet {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *(const uint4*)p);
}

__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0)); }

__device__ __forceinline__ float bf(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Block-wide FP64 sum of one value per thread into out[blockIdx.x].
__device__ __forceinline__ void block_sum(double a, double* out) {
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// ---------------------------------------------------------------------------
// Convolution forward (and backward-data), Cout = 16 * NB.  Cin and the channel
// strides are multiples of 8 so that every lane's 8 K-values are one 16-byte load of
// consecutive channels of one tap.
// ---------------------------------------------------------------------------
template <int NB>
__global__ __launch_bounds__(256) void conv3_kernel(const uint16_t* __restrict__ in, int cs_in,
                                                    int Cin, const uint16_t* __restrict__ W,
                                                    const float* __restrict__ bias,
                                                    uint16_t* __restrict__ out, int cs_out,
                                                    int relu, Grid3 g) {
  const long long V = (long long)g.Y * g.X * g.Z;
  const int lane = threadIdx.x & 63, row = lane & 15, grp = lane >> 4;
  const long long v0 = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
  if (v0 >= V) return;
  const long long XZ = (long long)g.X * g.Z;
  long long vm[2];
  int vy[2], vx[2], vz[2];
  bool vin[2];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb) {
    vm[mb] = v0 + 16 * mb + row;
    vin[mb] = vm[mb] < V;
    const long long t = vm[mb] / g.Z;
    vz[mb] = (int)(vm[mb] - t * g.Z);
    vy[mb] = (int)(t / g.X);
    vx[mb] = (int)(t - (long long)vy[mb] * g.X);
  }
  f32x4 acc[2][NB];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  constexpr int Cout = 16 * NB;
  const int K = 27 * Cin;
  int tap = 0, ci = 8 * grp;
  while (ci >= Cin) {
    ci -= Cin;
    ++tap;
  }
  for (int k0 = 0; k0 < K; k0 += 32) {
    const bool kv = tap < 27;
    const int dy = tap / 9 - 1, dx = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
    bf16x8 a[2], b[NB];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
      const int yy = vy[mb] + dy, xx = vx[mb] + dx, zz = vz[mb] + dz;
      const bool ok = kv && vin[mb] && yy >= 0 && yy < g.Y && xx >= 0 && xx < g.X && zz >= 0 &&
                      zz < g.Z;
      a[mb] = ok ? load8(in + (vm[mb] + dy * XZ + dx * g.Z + dz) * cs_in + ci) : zero8();
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
      b[nb] = kv ? load8(W + ((size_t)tap * Cout + 16 * nb + row) * Cin + ci) : zero8();
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = mfma(a[mb], b[nb], acc[mb][nb]);
    ci += 32;
    while (ci >= Cin) {
      ci -= Cin;
      ++tap;
    }
  }
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int co = 16 * nb + row;
      const float bb = bias ? bias[co] : 0.f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long m = v0 + 16 * mb + 4 * grp + r;
        if (m < V) {
          float val = acc[mb][nb][r] + bb;
          if (relu) val = fmaxf(val, 0.f);
          out[m * cs_out + co] = to_bf16(val);
        }
      }
    }
}

// ---------------------------------------------------------------------------
// The same convolution with its input staged in LDS.  A workgroup computes a box of
// kTY x kTX x kTZ = 4 x 4 x 16 output voxels (wave w: line y0 + w, whose four x lines are
// its four 16-row MFMA blocks); the box's input halo, 6 x 6 x 18 voxels with zeros
// outside the grid, is staged CC channels at a time, so every input element is read
// from memory about 2.5 times per output instead of 27.  K runs over the channel
// chunks, and within a chunk over k = tap * CC + c as in conv3_kernel.
// ---------------------------------------------------------------------------
constexpr int kTY = 4, kTX = 4, kTZ = 16;
constexpr int kHX = kTX + 2, kHZ = kTZ + 2;
constexpr int kHV = (kTY + 2) * kHX * kHZ;  // voxels of the halo box

template <int NB, int CC>
__global__ __launch_bounds__(256) void conv3_lds_kernel(
    const uint16_t* __restrict__ in, int cs_in, int Cin, const uint16_t* __restrict__ W,
    const float* __restrict__ bias, uint16_t* __restrict__ out, int cs_out, int relu, Grid3 g,
    int tiles_x, int tiles_z) {
  constexpr int Cout = 16 * NB;
  constexpr int VS = CC + 8;  // channels per staged voxel, padded against bank conflicts
  constexpr int U = CC / 8;   // 16-byte units per staged voxel
  __shared__ __align__(16) uint16_t halo[kHV * VS];
  const int lane = threadIdx.x & 63, row = lane & 15, grp = lane >> 4;
  const int wv = threadIdx.x >> 6;
  int t = xcd_logical_block(blockIdx.x, gridDim.x);
  const int tz = t % tiles_z;
  t /= tiles_z;
  const int tx = t % tiles_x, ty = t / tiles_x;
  const int y0 = ty * kTY, x0 = tx * kTX, z0 = tz * kTZ;
  f32x4 acc[4][NB];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int c0 = 0; c0 < Cin; c0 += CC) {
    for (int i = threadIdx.x; i < kHV * U; i += 256) {
      const int hv = i / U, u = i - hv * U;
      const int hz = hv % kHZ, hxy = hv / kHZ;
      const int hx = hxy % kHX, hy = hxy / kHX;
      const int yy = y0 - 1 + hy, xx = x0 - 1 + hx, zz = z0 - 1 + hz;
      uint4 val = make_uint4(0, 0, 0, 0);
      if (yy >= 0 && yy < g.Y && xx >= 0 && xx < g.X && zz >= 0 && zz < g.Z)
        val = *(const uint4*)(in + (((long long)yy * g.X + xx) * g.Z + zz) * cs_in + c0 + 8 * u);
      *(uint4*)(halo + hv * VS + 8 * u) = val;
    }
    __syncthreads();
#pragma unroll 2
    for (int k0 = 0; k0 < 27 * CC; k0 += 32) {
      const int kg = k0 + 8 * grp;
      const int tap0 = kg / CC, c = kg - tap0 * CC;
      const bool kv = tap0 < 27;  // the last step's padding taps add zeros
      const int tap = kv ? tap0 : 26;  // ... at addresses inside the arrays
      const int dy = tap / 9, dx = (tap / 3) % 3, dz = tap % 3;  // offsets into the halo box
      bf16x8 b[NB];
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
        b[nb] = kv ? load8(W + ((size_t)tap * Cout + 16 * nb + row) * Cin + c0 + c) : zero8();
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const int hv = ((wv + dy) * kHX + (mb + dx)) * kHZ + (row + dz);
        const bf16x8 a =
            kv ? __builtin_bit_cast(bf16x8, *(const uint4*)(halo + hv * VS + c)) : zero8();
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = mfma(a, b[nb], acc[mb][nb]);
      }
    }
    __syncthreads();
  }
  const int yy = y0 + wv;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    const int xx = x0 + mb;
    if (yy >= g.Y || xx >= g.X) continue;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int co = 16 * nb + row;
      const float bb = bias ? bias[co] : 0.f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int zz = z0 + 4 * grp + r;
        if (zz < g.Z) {
          float val = acc[mb][nb][r] + bb;
          if (relu) val = fmaxf(val, 0.f);
          out[(((long long)yy * g.X + xx) * g.Z + zz) * cs_out + co] = to_bf16(val);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradient: part[s][tap][co][ci] = sum over the voxels of split s of
// dY[v][co] * X[v + off(tap)][ci]; for tap 0 also bpart[s][co] = sum dY[v][co].
// MFMA with rows co (16 * NBO), columns ci (16 * NBI, ci >= Cin masked) and K = 32
// voxels: lane group l / 16 takes the 8 voxels 8 (l / 16) .. + 7 of the chunk, which
// share (y, x) because Z is a multiple of 8.  One wave per (tap, split).
// ---------------------------------------------------------------------------
template <int NBO, int NBI>
__global__ __launch_bounds__(256) void wgrad_kernel(const uint16_t* __restrict__ X, int cs_x,
                                                    int Cin, const uint16_t* __restrict__ dY,
                                                    int cs_dy, Grid3 g, int nsplit,
                                                    long long chunk, float* __restrict__ part,
                                                    float* __restrict__ bpart) {
  constexpr int Cout = 16 * NBO;
  const long long V = (long long)g.Y * g.X * g.Z;
  const int lane = threadIdx.x & 63, row = lane & 15, grp = lane >> 4;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= 27 * nsplit) return;
  const int tap = w / nsplit, s = w - tap * nsplit;
  const int dy = tap / 9 - 1, dx = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
  const long long XZ = (long long)g.X * g.Z;
  const long long vb = (long long)s * chunk;
  const long long ve = vb + chunk < V ? vb + chunk : V;
  f32x4 acc[NBO][NBI];
#pragma unroll
  for (int i = 0; i < NBO; ++i)
#pragma unroll
    for (int j = 0; j < NBI; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum[NBO];
#pragma unroll
  for (int i = 0; i < NBO; ++i) bsum[i] = 0.f;
  for (long long vc = vb; vc < ve; vc += 32) {
    const long long v8 = vc + 8 * grp;  // first of this lane's 8 voxels (z multiple of 8)
    const long long t = v8 / g.Z;
    const int z0 = (int)(v8 - t * g.Z);
    const int y = (int)(t / g.X), x = (int)(t - (long long)y * g.X);
    const bool inplane = v8 < ve && y + dy >= 0 && y + dy < g.Y && x + dx >= 0 && x + dx < g.X;
    bf16x8 a[NBO], b[NBI];
#pragma unroll
    for (int i = 0; i < NBO; ++i) {
      uint16_t h[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        h[j] = (v8 + j < ve) ? dY[(v8 + j) * cs_dy + 16 * i + row] : (uint16_t)0;
#pragma unroll
      for (int j = 0; j < 8; ++j) bsum[i] += bf(h[j]);
      a[i] = __builtin_bit_cast(bf16x8, h);
    }
#pragma unroll
    for (int jb = 0; jb < NBI; ++jb) {
      const int ci = 16 * jb + row;
      uint16_t h[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int zz = z0 + j + dz;
        const bool ok = inplane && ci < Cin && zz >= 0 && zz < g.Z;
        h[j] = ok ? X[(v8 + j + dy * XZ + dx * g.Z + dz) * cs_x + ci] : (uint16_t)0;
      }
      b[jb] = __builtin_bit_cast(bf16x8, h);
    }
#pragma unroll
    for (int i = 0; i < NBO; ++i)
#pragma unroll
      for (int jb = 0; jb < NBI; ++jb) acc[i][jb] = mfma(a[i], b[jb], acc[i][jb]);
  }
  float* P = part + ((size_t)s * 27 + tap) * Cout * Cin;
#pragma unroll
  for (int i = 0; i < NBO; ++i)
#pragma unroll
    for (int jb = 0; jb < NBI; ++jb) {
      const int ci = 16 * jb + row;
      if (ci < Cin) {
#pragma unroll
        for (int r = 0; r < 4; ++r) P[(size_t)(16 * i + 4 * grp + r) * Cin + ci] = acc[i][jb][r];
      }
    }
  if (tap == 0) {
#pragma unroll
    for (int i = 0; i < NBO; ++i) {
      float v = bsum[i];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      if (grp == 0) bpart[(size_t)s * Cout + 16 * i + row] = v;
    }
  }
}

// out[i] = sum_s part[s * n + i]  (fixed order)
__global__ void reduce_splits_kernel(const float* __restrict__ part, int nsplit, size_t n,
                                     float* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float a = 0.f;
  for (int s = 0; s < nsplit; ++s) a += part[(size_t)s * n + i];
  out[i] = a;
}

// Master weights [27][Cout][Cin] (FP32) -> forward BF16 copy (same layout) and the
// backward-data copy Wb[tap][ci][co] = W[26 - tap][co][ci].
__global__ void convert_weights_kernel(const float* __restrict__ Wm, int Cin, int Cout,
                                       uint16_t* __restrict__ Wf, uint16_t* __restrict__ Wb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 27 * Cin * Cout) return;
  const int tap = i / (Cin * Cout), rem = i - tap * Cin * Cout;
  const int co = rem / Cin, ci = rem - co * Cin;
  const uint16_t h = to_bf16(Wm[i]);
  Wf[i] = h;
  Wb[((size_t)(26 - tap) * Cin + ci) * Cout + co] = h;
}

// Network input: channels (x * s0, d * s1, 0 ...) of an 8-channel tensor.
__global__ void pack_input_kernel(const float* __restrict__ x, const float* __restrict__ d,
                                  float s0, float s1, long long V, uint16_t* __restrict__ out) {
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < V;
       v += (long long)gridDim.x * blockDim.x) {
    uint16_t h[8] = {to_bf16(x[v] * s0), to_bf16(d[v] * s1), 0, 0, 0, 0, 0, 0};
    *(uint4*)(out + v * 8) = __builtin_bit_cast(uint4, h);
  }
}

// 2x2x2 average pooling: out[p][c] = mean of the 8 children of p.
__global__ void pool_kernel(const uint16_t* __restrict__ in, int cs_in, int C, Grid3 gf,
                            uint16_t* __restrict__ out, int cs_out) {
  const int Yc = gf.Y / 2, Xc = gf.X / 2, Zc = gf.Z / 2;
  const long long n = (long long)Yc * Xc * Zc * C;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long p = i / C;
    const int c = (int)(i - p * C);
    const long long t = p / Zc;
    const int z = (int)(p - t * Zc), y = (int)(t / Xc), x = (int)(t - (long long)y * Xc);
    float s = 0.f;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const long long vf = ((long long)(2 * y + (o >> 2)) * gf.X + 2 * x + ((o >> 1) & 1)) * gf.Z +
                           2 * z + (o & 1);
      s += bf(in[vf * cs_in + c]);
    }
    out[p * cs_out + c] = to_bf16(0.125f * s);
  }
}

// Adjoint of pool_kernel, accumulated: din[child][c] += dout[p][c] / 8.
__global__ void pool_bwd_kernel(const uint16_t* __restrict__ dout, int cs_out, int C, Grid3 gf,
                                uint16_t* __restrict__ din, int cs_in) {
  const long long n = (long long)gf.Y * gf.X * gf.Z * C;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long v = i / C;
    const int c = (int)(i - v * C);
    const long long t = v / gf.Z;
    const int z = (int)(v - t * gf.Z), y = (int)(t / gf.X), x = (int)(t - (long long)y * gf.X);
    const long long p = ((long long)(y / 2) * (gf.X / 2) + x / 2) * (gf.Z / 2) + z / 2;
    const size_t o = (size_t)v * cs_in + c;
    din[o] = to_bf16(bf(din[o]) + 0.125f * bf(dout[p * cs_out + c]));
  }
}

// Nearest-neighbour 2x upsampling: out[child][c] = in[parent][c] (gf: the fine grid).
__global__ void upsample_kernel(const uint16_t* __restrict__ in, int cs_in, int C, Grid3 gf,
                                uint16_t* __restrict__ out, int cs_out) {
  const long long n = (long long)gf.Y * gf.X * gf.Z * C;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long v = i / C;
    const int c = (int)(i - v * C);
    const long long t = v / gf.Z;
    const int z = (int)(v - t * gf.Z), y = (int)(t / gf.X), x = (int)(t - (long long)y * gf.X);
    const long long p = ((long long)(y / 2) * (gf.X / 2) + x / 2) * (gf.Z / 2) + z / 2;
    out[(size_t)v * cs_out + c] = in[p * cs_in + c];
  }
}

// upsample_kernel with one workgroup per fine (y, x) line and 8 channels (16 bytes) per
// copy: channel counts, strides and both pointers 16-byte aligned.
__global__ __launch_bounds__(256) void upsample_line_kernel(const uint16_t* __restrict__ in,
                                                            int cs_in, int C, Grid3 gf,
                                                            uint16_t* __restrict__ out,
                                                            int cs_out) {
  const int line = blockIdx.x;  // y * X + x of the fine grid
  const int y = line / gf.X, x = line - y * gf.X;
  const long long pline = (long long)(y / 2) * (gf.X / 2) + x / 2;
  const int U = C / 8;
  const uint4* src = (const uint4*)in;
  uint4* dst = (uint4*)out;
  for (int i = threadIdx.x; i < gf.Z * U; i += blockDim.x) {
    const int z = i / U, u = i - z * U;
    const long long p = pline * (gf.Z / 2) + z / 2;
    const long long v = (long long)line * gf.Z + z;
    dst[(v * cs_out) / 8 + u] = src[(p * cs_in) / 8 + u];
  }
}

// Adjoint of upsample_kernel: din[p][c] = sum of dout over the 8 children.
__global__ void upsample_bwd_kernel(const uint16_t* __restrict__ dout, int cs_out, int C,
                                    Grid3 gf, uint16_t* __restrict__ din, int cs_in) {
  const int Yc = gf.Y / 2, Xc = gf.X / 2, Zc = gf.Z / 2;
  const long long n = (long long)Yc * Xc * Zc * C;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long p = i / C;
    const int c = (int)(i - p * C);
    const long long t = p / Zc;
    const int z = (int)(p - t * Zc), y = (int)(t / Xc), x = (int)(t - (long long)y * Xc);
    float s = 0.f;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const long long vf = ((long long)(2 * y + (o >> 2)) * gf.X + 2 * x + ((o >> 1) & 1)) * gf.Z +
                           2 * z + (o & 1);
      s += bf(dout[vf * cs_out + c]);
    }
    din[p * cs_in + c] = to_bf16(s);
  }
}

// ReLU backward: dst[v][c] = (y[v][c] > 0) ? g[v][c] : 0 (dst may alias g).
__global__ void relu_bwd_kernel(const uint16_t* y, int cs_y, const uint16_t* gsrc, int cs_g,
                                int C, long long V, uint16_t* dst, int cs_d) {
  const long long n = V * C;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long v = i / C;
    const int c = (int)(i - v * C);
    const uint16_t gv = gsrc[v * cs_g + c];
    dst[v * cs_d + c] = bf(y[v * cs_y + c]) > 0.f ? gv : (uint16_t)0;
  }
}

// Head: h[v] = sum_c w[c] a[v][c] + b (FP32 output), 16 channels.
__global__ void head_kernel(const uint16_t* __restrict__ a, const float* __restrict__ wb,
                            long long V, float scale, float* __restrict__ h) {
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < V;
       v += (long long)gridDim.x * blockDim.x) {
    const uint4 q0 = *(const uint4*)(a + v * 16), q1 = *(const uint4*)(a + v * 16 + 8);
    const uint16_t* e0 = (const uint16_t*)&q0;
    const uint16_t* e1 = (const uint16_t*)&q1;
    float s = wb[16];
#pragma unroll
    for (int c = 0; c < 8; ++c) s = fmaf(wb[c], bf(e0[c]), s);
#pragma unroll
    for (int c = 0; c < 8; ++c) s = fmaf(wb[8 + c], bf(e1[c]), s);
    h[v] = s * scale;
  }
}

// Loss L = sum_v e_v^2 / (2 V) with e = h - t * tscale: dh = e / V, and part[block]
// the block's partial sum of e^2 (FP64).
__global__ __launch_bounds__(256) void mse_kernel(const float* __restrict__ h,
                                                  const float* __restrict__ t, float tscale,
                                                  long long V, float* __restrict__ dh,
                                                  double* __restrict__ part) {
  double s = 0.0;
  const float invV = 1.0f / (float)V;
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < V;
       v += (long long)gridDim.x * blockDim.x) {
    const float e = h[v] - t[v] * tscale;
    dh[v] = e * invV;
    s += (double)e * (double)e;
  }
  block_sum(s, part);
}

// Head backward with the ReLU of its input: da[v][c] = (a > 0) ? dh[v] w[c] : 0, and
// per-block partials of dw[c] = sum dh a[v][c] (c < 16) and db = sum dh.
__global__ __launch_bounds__(256) void head_bwd_kernel(const uint16_t* __restrict__ a,
                                                       const float* __restrict__ wb,
                                                       const float* __restrict__ dh, long long V,
                                                       uint16_t* __restrict__ da,
                                                       double* __restrict__ part) {
  double acc[17];
#pragma unroll
  for (int c = 0; c < 17; ++c) acc[c] = 0.0;
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < V;
       v += (long long)gridDim.x * blockDim.x) {
    const float g = dh[v];
    uint16_t out[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const float av = bf(a[v * 16 + c]);
      acc[c] += (double)g * av;
      out[c] = av > 0.f ? to_bf16(g * wb[c]) : (uint16_t)0;
    }
    acc[16] += g;
    uint16_t o0[8], o1[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      o0[c] = out[c];
      o1[c] = out[8 + c];
    }
    *(uint4*)(da + v * 16) = __builtin_bit_cast(uint4, o0);
    *(uint4*)(da + v * 16 + 8) = __builtin_bit_cast(uint4, o1);
  }
  __shared__ double red[17][4];
#pragma unroll
  for (int c = 0; c < 17; ++c) {
    double x = acc[c];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63) == 0) red[c][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < 17) {
    const int c = threadIdx.x;
    part[(size_t)blockIdx.x * 17 + c] = red[c][0] + red[c][1] + red[c][2] + red[c][3];
  }
}

// out[i] = sum over blocks b of part[b * n + i] (FP64 sum, FP32 result), i < n <= 64.
__global__ void reduce_dpart_kernel(const double* __restrict__ part, int nb, int n,
                                    float* __restrict__ out) {
  const int i = threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += part[(size_t)b * n + i];
  out[i] = (float)s;
}

// Adam on FP32 master parameters.
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, size_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

int grid1(long long n) {
  long long g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

int check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CTR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return CTR_OK;
}

}  // namespace

int conv3(const uint16_t* in, int cs_in, int Cin, const uint16_t* W, const float* bias,
          uint16_t* out, int cs_out, int Cout, bool relu, Grid3 g, hipStream_t st) {
  if (Cin % 8 || cs_in % 8 || cs_out % 8 || Cout % 16)
    return fail(CTR_ERR_ARG, "conv3: channel counts and strides must be multiples of 8");
  const int NB = Cout / 16;
  if ((Cin == 8 || Cin % 16 == 0) && NB <= 4) {
    // LDS-staged input halo: 16 channels per stage (8 for an 8-channel input)
    const int tiles_x = (g.X + kTX - 1) / kTX, tiles_z = (g.Z + kTZ - 1) / kTZ;
    const long long tiles = (long long)((g.Y + kTY - 1) / kTY) * tiles_x * tiles_z;
    const dim3 tgrid((unsigned)tiles);
#define CVL(NBV, CCV)                                                                           \
  hipLaunchKernelGGL((conv3_lds_kernel<NBV, CCV>), tgrid, dim3(256), 0, st, in, cs_in, Cin, W, \
                     bias, out, cs_out, (int)relu, g, tiles_x, tiles_z)
    if (Cin == 8) {
      switch (NB) {
        case 1: CVL(1, 8); break;
        case 2: CVL(2, 8); break;
        case 3: CVL(3, 8); break;
        default: CVL(4, 8); break;
      }
    } else {
      switch (NB) {
        case 1: CVL(1, 16); break;
        case 2: CVL(2, 16); break;
        case 3: CVL(3, 16); break;
        default: CVL(4, 16); break;
      }
    }
#undef CVL
    return check("conv3_lds_kernel");
  }
  const long long V = (long long)g.Y * g.X * g.Z;
  const long long waves = (V + 31) / 32;
  dim3 grid((unsigned)((waves + 3) / 4));
#define CV(NB)                                                                                   \
  hipLaunchKernelGGL(conv3_kernel<NB>, grid, dim3(256), 0, st, in, cs_in, Cin, W, bias, out, \
                     cs_out, (int)relu, g)
  switch (Cout / 16) {
    case 1: CV(1); break;
    case 2: CV(2); break;
    case 3: CV(3); break;
    case 4: CV(4); break;
    case 6: CV(6); break;
    default: return fail(CTR_ERR_UNSUPPORTED, "conv3: output channel count not instantiated");
  }
#undef CV
  return check("conv3_kernel");
}

int wgrad_splits(Grid3 g) {
  const long long V = (long long)g.Y * g.X * g.Z;
  long long s = V / 2048;  // at least 2048 voxels per wave
  if (s < 1) s = 1;
  if (s > 64) s = 64;
  return (int)s;
}

int wgrad(const uint16_t* X, int cs_x, int Cin, const uint16_t* dY, int cs_dy, int Cout, Grid3 g,
          float* part, float* bpart, float* dW, float* db, hipStream_t st) {
  if (g.Z % 8) return fail(CTR_ERR_ARG, "wgrad: Z must be a multiple of 8");
  const long long V = (long long)g.Y * g.X * g.Z;
  const int ns = wgrad_splits(g);
  long long chunk = (V + ns - 1) / ns;
  chunk = (chunk + 31) / 32 * 32;
  dim3 grid((27 * ns + 3) / 4);
  const int nbo = Cout / 16, nbi = (Cin + 15) / 16;
#define WG(O, I)                                                                               \
  hipLaunchKernelGGL((wgrad_kernel<O, I>), grid, dim3(256), 0, st, X, cs_x, Cin, dY, cs_dy, g, \
                     ns, chunk, part, bpart)
  if (nbo == 1 && nbi == 1) WG(1, 1);
  else if (nbo == 1 && nbi == 3) WG(1, 3);
  else if (nbo == 2 && nbi == 1) WG(2, 1);
  else if (nbo == 2 && nbi == 2) WG(2, 2);
  else if (nbo == 2 && nbi == 6) WG(2, 6);
  else if (nbo == 4 && nbi == 2) WG(4, 2);
  else if (nbo == 4 && nbi == 4) WG(4, 4);
  else return fail(CTR_ERR_UNSUPPORTED, "wgrad: channel combination not instantiated");
#undef WG
  int rc = check("wgrad_kernel");
  if (rc) return rc;
  const size_t n = (size_t)27 * Cout * Cin;
  hipLaunchKernelGGL(reduce_splits_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     part, ns, n, dW);
  hipLaunchKernelGGL(reduce_splits_kernel, dim3((unsigned)((Cout + 255) / 256)), dim3(256), 0, st,
                     bpart, ns, (size_t)Cout, db);
  return check("reduce_splits_kernel");
}

int convert_weights(const float* Wm, int Cin, int Cout, uint16_t* Wf, uint16_t* Wb,
                    hipStream_t st) {
  const int n = 27 * Cin * Cout;
  hipLaunchKernelGGL(convert_weights_kernel, dim3((n + 255) / 256), dim3(256), 0, st, Wm, Cin,
                     Cout, Wf, Wb);
  return check("convert_weights_kernel");
}

int pack_input(const float* x, const float* d, float s0, float s1, long long V, uint16_t* out,
               hipStream_t st) {
  hipLaunchKernelGGL(pack_input_kernel, dim3(grid1(V)), dim3(256), 0, st, x, d, s0, s1, V, out);
  return check("pack_input_kernel");
}

int pool(const uint16_t* in, int cs_in, int C, Grid3 gf, uint16_t* out, int cs_out,
         hipStream_t st) {
  const long long n = (long long)gf.Y * gf.X * gf.Z / 8 * C;
  hipLaunchKernelGGL(pool_kernel, dim3(grid1(n)), dim3(256), 0, st, in, cs_in, C, gf, out, cs_out);
  return check("pool_kernel");
}

int pool_bwd(const uint16_t* dout, int cs_out, int C, Grid3 gf, uint16_t* din, int cs_in,
             hipStream_t st) {
  const long long n = (long long)gf.Y * gf.X * gf.Z * C;
  hipLaunchKernelGGL(pool_bwd_kernel, dim3(grid1(n)), dim3(256), 0, st, dout, cs_out, C, gf, din,
                     cs_in);
  return check("pool_bwd_kernel");
}

int upsample(const uint16_t* in, int cs_in, int C, Grid3 gf, uint16_t* out, int cs_out,
             hipStream_t st) {
  if (C % 8 == 0 && cs_in % 8 == 0 && cs_out % 8 == 0 && ((uintptr_t)in & 15) == 0 &&
      ((uintptr_t)out & 15) == 0 && (long long)gf.Y * gf.X <= 0x7fffffffLL) {
    hipLaunchKernelGGL(upsample_line_kernel, dim3((unsigned)(gf.Y * gf.X)), dim3(256), 0, st,
                       in, cs_in, C, gf, out, cs_out);
    return check("upsample_line_kernel");
  }
  const long long n = (long long)gf.Y * gf.X * gf.Z * C;
  hipLaunchKernelGGL(upsample_kernel, dim3(grid1(n)), dim3(256), 0, st, in, cs_in, C, gf, out,
                     cs_out);
  return check("upsample_kernel");
}

int upsample_bwd(const uint16_t* dout, int cs_out, int C, Grid3 gf, uint16_t* din, int cs_in,
                 hipStream_t st) {
  const long long n = (long long)gf.Y * gf.X * gf.Z / 8 * C;
  hipLaunchKernelGGL(upsample_bwd_kernel, dim3(grid1(n)), dim3(256), 0, st, dout, cs_out, C, gf,
                     din, cs_in);
  return check("upsample_bwd_kernel");
}

int relu_bwd(const uint16_t* y, int cs_y, const uint16_t* g, int cs_g, int C, long long V,
             uint16_t* dst, int cs_d, hipStream_t st) {
  hipLaunchKernelGGL(relu_bwd_kernel, dim3(grid1(V * C)), dim3(256), 0, st, y, cs_y, g, cs_g, C,
                     V, dst, cs_d);
  return check("relu_bwd_kernel");
}

int head(const uint16_t* a, const float* wb, long long V, float scale, float* h, hipStream_t st) {
  hipLaunchKernelGGL(head_kernel, dim3(grid1(V)), dim3(256), 0, st, a, wb, V, scale, h);
  return check("head_kernel");
}

int mse(const float* h, const float* t, float tscale, long long V, float* dh, double* part,
        int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(mse_kernel, dim3(nblocks), dim3(256), 0, st, h, t, tscale, V, dh, part);
  return check("mse_kernel");
}

int head_bwd(const uint16_t* a, const float* wb, const float* dh, long long V, uint16_t* da,
             double* part, int nblocks, float* dwb, hipStream_t st) {
  hipLaunchKernelGGL(head_bwd_kernel, dim3(nblocks), dim3(256), 0, st, a, wb, dh, V, da, part);
  int rc = check("head_bwd_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(reduce_dpart_kernel, dim3(1), dim3(64), 0, st, part, nblocks, 17, dwb);
  return check("reduce_dpart_kernel");
}

int adam(float* p, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
         float eps, int step, hipStream_t st) {
  const float c1 = 1.f / (1.f - powf(b1, (float)step));
  const float c2 = 1.f / (1.f - powf(b2, (float)step));
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, g, m, v,
                     n, lr, b1, b2, eps, c1, c2);
  return check("adam_kernel");
}

}  // namespace unet
}  // namespace ctr
