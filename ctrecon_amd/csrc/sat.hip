// Integral-image (summed-area table) method: compensated FP32 integral images built
// with wavefront-level scans, and the forward / reference back projector that sample
// them with clamped bilinear interpolation.
//
// This is the reference's own formulation (DD3_GPU.cu:33-50, 101-197 volume SATs,
// 1241-1323 sinogram SATs, 844-990 forward, 1397-1475 back), with two changes:
//  * prefix sums are double-float (hi + lo FP32 pairs accumulated with TwoSum, each
//    texel stored as the pair (hi, lo)), so the difference of four nearby corners
//    keeps its low-order bits; whole-plane FP32 integral images lose ~eps * 1e5 per
//    corner at clinical sizes;
//  * texture fetches are software bilinear samples with exact weights and clamp
//    addressing (sample at t - 0.5, clamp each texel index), evaluated in FP64.
// The direct-footprint kernels in kernels.hip compute the same operator faster;
// this method is selected with ctr_plan_set_method(plan, CTR_METHOD_SAT).
#include <algorithm>

#include "ctr_internal.hpp"
#include "device_common.hpp"

namespace ctr {
namespace {

// Source of the value at padded position (slice s, row r', column c'); r' = 0 and
// c' = 0 are the zero border (DD3_GPU.cu:45-47, 1250-1251).
template <typename T>
struct VolumeRowsX {  // slices x, rows y', columns z' (the ZYX table)
  const T* vol;
  const float* maskf;
  int XN, YN, ZN;
  __device__ float operator()(int s, int r, int c) const {
    const int col = (r - 1) * XN + s;
    return ld(vol, (size_t)col * ZN + (c - 1)) * maskf[col];
  }
};
template <typename T>
struct VolumeRowsY {  // slices y, rows x', columns z' (the ZXY table)
  const T* vol;
  const float* maskf;
  int XN, YN, ZN;
  __device__ float operator()(int s, int r, int c) const {
    const int col = s * XN + (r - 1);
    return ld(vol, (size_t)col * ZN + (c - 1)) * maskf[col];
  }
};
template <typename T>
struct SinoRows {  // slices k (views of the launch), rows u', columns v'
  const T* prj;
  int DNU, DNV;
  __device__ float operator()(int s, int r, int c) const {
    return ld(prj, ((size_t)s * DNU + (r - 1)) * DNV + (c - 1));
  }
};

// Pass 1: inclusive prefix along the contiguous axis (columns c' of one row) with a
// wave-level scan, 64 columns at a time with a carried double-float total.
// One wave per (slice, row); rows 0 (the border) are written as zeros.
template <typename Src>
__global__ __launch_bounds__(kBlock) void sat_rows_kernel(Src src, int NSl, int R, int W,
                                                          float2* __restrict__ hl) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (row >= NSl * (R + 1)) return;
  const int s = row / (R + 1), r = row - s * (R + 1);
  const size_t base = (size_t)row * W;
  DF carry{0.f, 0.f};
  for (int c0 = 0; c0 < W; c0 += 64) {
    const int c = c0 + lane;
    DF v{(r >= 1 && c >= 1 && c < W) ? src(s, r, c) : 0.f, 0.f};
    for (int off = 1; off < 64; off <<= 1) {
      DF o;
      o.hi = __shfl_up(v.hi, off, 64);
      o.lo = __shfl_up(v.lo, off, 64);
      if (lane >= off) v = df_add(o, v);
    }
    v = df_add(carry, v);
    if (c < W) hl[base + c] = make_float2(v.hi, v.lo);
    carry.hi = __shfl(v.hi, 63, 64);
    carry.lo = __shfl(v.lo, 63, 64);
  }
}

// Pass 2: inclusive prefix along the rows of each slice (stride W); one lane per
// (slice, column), lanes across columns so every access is coalesced.
__global__ __launch_bounds__(kBlock) void sat_cols_kernel(int R, int W, float2* __restrict__ hl) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  const int s = blockIdx.y;
  if (c >= W) return;
  size_t o = ((size_t)s * (R + 1)) * W + c;
  const float2 v0 = hl[o];
  DF acc{v0.x, v0.y};
  for (int r = 1; r <= R; ++r) {
    o += W;
    const float2 v = hl[o];
    acc = df_add(acc, DF{v.x, v.y});
    hl[o] = make_float2(acc.hi, acc.lo);
  }
}

// Clamped bilinear sample of the double-float table plane at unnormalised texture
// coordinates (a along the W axis, b along the H axis), as cudaFilterModeLinear +
// cudaAddressModeClamp define it, with exact weights.
__device__ __forceinline__ double sat_sample(const BufF32& HL, int base, int W, int H, double a,
                                             double b) {
  a = fmin(fmax(a - 0.5, -1.0), (double)W);
  b = fmin(fmax(b - 0.5, -1.0), (double)H);
  const double fa = floor(a), fb = floor(b);
  const int i = (int)fa, j = (int)fb;
  const double al = a - fa, be = b - fb;
  const int i0 = min(max(i, 0), W - 1), i1 = min(max(i + 1, 0), W - 1);
  const int j0 = min(max(j, 0), H - 1), j1 = min(max(j + 1, 0), H - 1);
  // texels i0, i1 of a row are among the adjacent pair q, q + 1 (q = i inside the row,
  // the end pair when clamped); texel t is (hi, lo) at floats 2t, 2t + 1, so a row's
  // pair is one run of four floats
  const int q = min(max(i, 0), W - 2);
  const bool s0 = i0 == q, s1 = i1 == q;
  const int r0 = 2 * (base + j0 * W + q), r1 = 2 * (base + j1 * W + q);
  const float h0x = HL[r0], l0x = HL[r0 + 1], h0y = HL[r0 + 2], l0y = HL[r0 + 3];
  const float h1x = HL[r1], l1x = HL[r1 + 1], h1y = HL[r1 + 2], l1y = HL[r1 + 3];
  const double v00 = (double)(s0 ? h0x : h0y) + (double)(s0 ? l0x : l0y);
  const double v01 = (double)(s1 ? h0x : h0y) + (double)(s1 ? l0x : l0y);
  const double v10 = (double)(s0 ? h1x : h1y) + (double)(s0 ? l1x : l1y);
  const double v11 = (double)(s1 ? h1x : h1y) + (double)(s1 ? l1x : l1y);
  return (1.0 - be) * ((1.0 - al) * v00 + al * v01) + be * ((1.0 - al) * v10 + al * v11);
}

// Forward projector on the integral images (DD3_GPU.cu:844-990): per slice the four
// corner samples of the footprint, divided by the footprint area.  TO: storage
// type of out (float, or uint16_t for BF16).
template <bool FUSED, typename TO>
__global__ __launch_bounds__(kBlock) void fp_sat_kernel(
    const float2* __restrict__ zyx, const float2* __restrict__ zxy, TO* __restrict__ out, const ViewP* __restrict__ views, const float2* __restrict__ bnd,
    const float4* __restrict__ cols, const float* __restrict__ Bz,
    const float* __restrict__ h2, Dims d, int vb, int vs, const float* __restrict__ b,
    const float* __restrict__ invRow, double* __restrict__ partials) {
  const int hw = blockIdx.y * gridDim.x + blockIdx.x;
  const int lb = xcd_logical_block(hw, gridDim.x * gridDim.y);
  const int k = lb / gridDim.x;
  const int p = vb + k * vs;
  const int ray = (lb - k * gridDim.x) * kBlock + threadIdx.x;
  double res = 0.0, bsq = 0.0;
  if (ray < d.DNU * d.DNV) {
    const int u = ray / d.DNV, v = ray - u * d.DNV;
    const ViewP vp = views[p];
    const float2 bl = bnd[(size_t)p * (d.DNU + 1) + u];
    const float2 br = bnd[(size_t)p * (d.DNU + 1) + u + 1];
    const float4 ct = cols[(size_t)p * d.DNU + u];
    const float Bv = Bz[v], Bv1 = Bz[v + 1];
    const int NS = vp.xmarch ? d.XN : d.YN;
    const int H = (vp.xmarch ? d.YN : d.XN) + 1, W = d.ZN + 1;
    const uint32_t n = (uint32_t)NS * H * W;
    const BufF32 HL(reinterpret_cast<const float*>(vp.xmarch ? zyx : zxy), 2 * n);
    double acc = 0.0;
    for (int ii = 0; ii < NS; ++ii) {
      const float fi = (float)ii;
      const double tl = fmaf(fi, bl.y, bl.x) + 1.0, tr = fmaf(fi, br.y, br.x) + 1.0;
      const float r = fmaf(fi, ct.y, ct.x);
      const double tD = fmaf(r, Bv, vp.az) + 1.0, tU = fmaf(r, Bv1, vp.az) + 1.0;
      // a footprint wholly outside the volume in z or in-plane clamps the samples of a
      // pair to the same texels: the slice adds exactly 0 and is skipped
      if (fmax(tU, tD) - 0.5 <= -1.0 || fmin(tU, tD) - 0.5 >= (double)W ||
          fmax(tl, tr) - 0.5 <= -1.0 || fmin(tl, tr) - 0.5 >= (double)H)
        continue;
      const int base = ii * H * W;
      // grouped as two differences so that samples clamped to the same texels
      // cancel exactly (a footprint outside the volume gives exactly 0)
      const double diff = (sat_sample(HL, base, W, H, tU, tr) -
                           sat_sample(HL, base, W, H, tD, tr)) -
                          (sat_sample(HL, base, W, H, tU, tl) -
                           sat_sample(HL, base, W, H, tD, tl));
      acc += diff / ((tr - tl) * (tU - tD));
    }
    const float val = (float)(acc * (double)(ct.w * __builtin_amdgcn_sqrtf(ct.z + h2[v])));
    const size_t o = ((size_t)k * d.DNU + u) * d.DNV + v;
    if constexpr (FUSED) {
      const size_t og = ((size_t)p * d.DNU + u) * d.DNV + v;
      const float bb = b[og];
      const float diffv = bb - val;
      st(out, o, invRow ? diffv * invRow[o] : diffv);
      res = (double)diffv * (double)diffv;
      bsq = (double)bb * (double)bb;
    } else {
      st(out, o, val);
    }
  }
  if constexpr (FUSED) {
    __shared__ double red[2][kWaves];
    for (int off = 32; off > 0; off >>= 1) {
      res += __shfl_xor(res, off, 64);
      bsq += __shfl_xor(bsq, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      red[0][threadIdx.x >> 6] = res;
      red[1][threadIdx.x >> 6] = bsq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double a = 0, c = 0;
      for (int i = 0; i < kWaves; ++i) {
        a += red[0][i];
        c += red[1][i];
      }
      partials[2 * hw] = a;
      partials[2 * hw + 1] = c;
    }
  }
}

// Reference back projector on the sinogram integral images (DD3_GPU.cu:1397-1475):
// one wave per voxel column, one voxel per lane.  UPDATE also refreshes the BF16
// shadow of vol when shadow is not NULL.
template <bool UPDATE>
__global__ __launch_bounds__(kBlock) void bp_sat_kernel(
    const float2* __restrict__ sat, float* __restrict__ vol,
    const ViewP* __restrict__ views, const double2* __restrict__ src64,
    const uint8_t* __restrict__ mask, Dims d, BpConsts c, BpConsts64 c64, int vb, int vs,
    int vc, const float* __restrict__ invCol, uint16_t* __restrict__ shadow) {
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * kWaves + (threadIdx.x >> 6);
  const int z = blockIdx.y * 64 + lane;
  if (col >= d.XN * d.YN || z >= d.ZN) return;
  const size_t o = (size_t)col * d.ZN + z;
  if (mask[col] != 1) {
    if constexpr (!UPDATE) vol[o] = 0.f;
    return;
  }
  const int y = col / d.XN, x = col - y * d.XN;
  const double Px = ((double)x - c64.ocx) * c64.dx;
  const double Py = ((double)y - c64.ocy) * c64.dx;
  const double Pz = ((double)z - (double)d.ocz) * (double)d.dz;
  const int W = d.DNV + 1, H = d.DNU + 1;
  const BufF32 HL(reinterpret_cast<const float*>(sat), 2u * (uint32_t)vc * W * H);
  double acc = 0.0;
  for (int k = 0; k < vc; ++k) {
    const int p = vb + k * vs;
    const ViewP vp = views[p];
    const double2 s64 = src64[p];
    const double dX = Px - s64.x, dY = Py - s64.y, dZ = Pz - (double)vp.sz;
    const double l2 = dX * dX + dY * dY;
    const double rl = rsqrt(l2);
    const double iU = (dZ + c.hz) * c.S * rl + c.cV + 1.0;
    const double iD = (dZ - c.hz) * c.S * rl + c.cV + 1.0;
    // a row shadow wholly below or above the detector clamps all four samples of a
    // column pair to the same texels: the view adds exactly 0 and is skipped
    if (fmax(iU, iD) - 0.5 <= -1.0 || fmin(iU, iD) - 0.5 >= (double)W) continue;
    const double alpha = asin((s64.y * Px - s64.x * Py) * c64.inv_sid * rl);
    const double w = vp.wx ? dX : dY;
    const double dA = w * (rl * rl) * (c64.dx * 0.5);
    const double cosVal = c64.dx / w * sqrt(l2 + dZ * dZ);
    const double iL = (alpha - dA) * c64.ib + c64.cU + 1.0;
    const double iR = (alpha + dA) * c64.ib + c64.cU + 1.0;
    const int base = k * W * H;
    // grouped as two differences so that a shadow outside the detector gives exactly
    // 0: the OS-SART column-sum guard (os_sart.m:38-39) relies on exact zeros
    acc += ((sat_sample(HL, base, W, H, iU, iR) - sat_sample(HL, base, W, H, iD, iR)) -
            (sat_sample(HL, base, W, H, iU, iL) - sat_sample(HL, base, W, H, iD, iL))) *
           cosVal;
  }
  if constexpr (UPDATE) {
    const float nx = fmaf((float)acc, invCol[o], vol[o]);
    vol[o] = nx;
    if (shadow) shadow[o] = to_bf16_dither(nx, o);
  } else {
    vol[o] = (float)acc;
  }
}

int check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CTR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return CTR_OK;
}

template <typename Src>
int build_sat(Src src, int NSl, int R, int W, float2* hl, hipStream_t st) {
  const int rows = NSl * (R + 1);
  hipLaunchKernelGGL(sat_rows_kernel<Src>, dim3((rows + kWaves - 1) / kWaves), dim3(kBlock), 0,
                     st, src, NSl, R, W, hl);
  int rc = check("sat_rows_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(sat_cols_kernel, dim3((W + kBlock - 1) / kBlock, NSl), dim3(kBlock), 0, st,
                     R, W, hl);
  return check("sat_cols_kernel");
}

}  // namespace

size_t sat_volume_floats(const Dims& d) {
  return (size_t)d.XN * (d.YN + 1) * (d.ZN + 1) + (size_t)d.YN * (d.XN + 1) * (d.ZN + 1);
}

size_t sat_sino_floats(const Dims& d, int vc) {
  return (size_t)vc * (d.DNU + 1) * (d.DNV + 1);
}

// Volume integral images into sat: ZYX then ZXY texels, each texel (hi, lo).
template <typename T>
int launch_volume_sat(const Plan& P, const T* vol, float* sat, hipStream_t st) {
  const Dims& d = P.d;
  const size_t nx = (size_t)d.XN * (d.YN + 1) * (d.ZN + 1);
  float2* hl = reinterpret_cast<float2*>(sat);
  int rc = build_sat(VolumeRowsX<T>{vol, P.maskf, d.XN, d.YN, d.ZN}, d.XN, d.YN, d.ZN + 1, hl,
                     st);
  if (rc) return rc;
  return build_sat(VolumeRowsY<T>{vol, P.maskf, d.XN, d.YN, d.ZN}, d.YN, d.XN, d.ZN + 1,
                   hl + nx, st);
}

template <typename TO>
int launch_fp_sat(const Plan& P, const float* sat, TO* out, int vb, int vs, int vc,
                  hipStream_t st, const float* b, const float* invRow, double* partials) {
  const Dims& d = P.d;
  const size_t nx = (size_t)d.XN * (d.YN + 1) * (d.ZN + 1);
  const float2* zyx = reinterpret_cast<const float2*>(sat);
  const float2* zxy = zyx + nx;
  dim3 grid((d.DNU * d.DNV + kBlock - 1) / kBlock, vc);
  if (b)
    hipLaunchKernelGGL((fp_sat_kernel<true, TO>), grid, dim3(kBlock), 0, st, zyx, zxy, out,
                       P.views, P.bnd, P.cols, P.Bz, P.h2, d, vb, vs, b, invRow, partials);
  else
    hipLaunchKernelGGL((fp_sat_kernel<false, TO>), grid, dim3(kBlock), 0, st, zyx, zxy, out,
                       P.views, P.bnd, P.cols, P.Bz, P.h2, d, vb, vs, nullptr, nullptr,
                       nullptr);
  return check("fp_sat_kernel");
}

template <typename T>
int launch_bp_sat(const Plan& P, const T* prj, float* sat, float* vol, int vb, int vs, int vc,
                  hipStream_t st, const float* invCol, int accumulate, uint16_t* shadow) {
  const Dims& d = P.d;
  float2* hl = reinterpret_cast<float2*>(sat);  // texels (hi, lo)
  int rc = build_sat(SinoRows<T>{prj, d.DNU, d.DNV}, vc, d.DNU, d.DNV + 1, hl, st);
  if (rc) return rc;
  dim3 grid((d.XN * d.YN + kWaves - 1) / kWaves, (d.ZN + 63) / 64);
  if (accumulate)
    hipLaunchKernelGGL(bp_sat_kernel<true>, grid, dim3(kBlock), 0, st, hl, vol, P.views,
                       P.src64, P.mask, d, P.bc, P.bc64, vb, vs, vc, invCol, shadow);
  else
    hipLaunchKernelGGL(bp_sat_kernel<false>, grid, dim3(kBlock), 0, st, hl, vol, P.views,
                       P.src64, P.mask, d, P.bc, P.bc64, vb, vs, vc, nullptr, nullptr);
  return check("bp_sat_kernel");
}

#define CTR_SAT_INST(T)                                                                       \
  template int launch_volume_sat<T>(const Plan&, const T*, float*, hipStream_t);              \
  template int launch_fp_sat<T>(const Plan&, const float*, T*, int, int, int, hipStream_t,    \
                                const float*, const float*, double*);                         \
  template int launch_bp_sat<T>(const Plan&, const T*, float*, float*, int, int, int,         \
                                hipStream_t, const float*, int, uint16_t*);
CTR_SAT_INST(float)
CTR_SAT_INST(uint16_t)
#undef CTR_SAT_INST

}  // namespace ctr
