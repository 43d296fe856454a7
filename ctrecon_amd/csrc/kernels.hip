// gfx950 kernels of the DD3 projector pair and the solver loops.
//
// Work mapping ("z-column waves"): every wave64 owns one column of 64 consecutive
// z-cells -- in the forward projector the 64 detector rows of one (view, column),
// in the back projectors 64 voxels along z of one (x, y) column.  All in-plane
// geometry is then the same for the wave and only the z part is per lane; reads
// of vol[...][z] and prj[...][v] are contiguous across the lanes.
//
// The reference samples 2-D summed-area tables through linear-filtered, clamped
// textures.  With exact weights that sample equals the integral of the
// piecewise-constant volume (sinogram) over the footprint rectangle clipped to
// the array, so the kernels integrate the footprint directly: the same operator,
// without the FP32 cancellation of whole-plane integral images and without the
// integral-image build pass.
#include <algorithm>
#include <type_traits>

#include "ctr_internal.hpp"
#include "device_common.hpp"

namespace ctr {
namespace {

// A lane's float as a wave-uniform value (v_readlane_b32 into a scalar register).
__device__ __forceinline__ float rdl(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Block-wide sum of two doubles into slot[0], slot[1].
__device__ __forceinline__ void block_sum2(double a, double b, double* slot) {
  __shared__ double red[2][kWaves];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = a;
    red[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0, sb = 0;
    for (int i = 0; i < kWaves; ++i) {
      sa += red[0][i];
      sb += red[1][i];
    }
    gp(slot)[0] = sa;
    gp(slot)[1] = sb;
  }
  __syncthreads();  // red[] is reused by the next work item of a persistent block
}

// ---------------------------------------------------------------------------
// Forward projector (mode 0 distance-driven), one thread per ray (p, u, v).
// Reference: DD3_gpu_proj_branchless_sat2d_ker, iterativeRecon/DD3_GPU.cu:844-990.
// Per slice ii the ray's footprint is [tl, tr] (in-plane) x [zl, zh] (z) in
// cell-centre index units; its mean over the footprint, summed over slices and
// scaled by dx/|dir| (K), is the ray sum.
// FUSED: out = (b - A x) * invRow (os_sart.m:24,30-31; b indexed by global view,
// invRow by the launch's compact view index, NULL for 1; out NULL: not written) and
// per-block sums of (b - A x)^2 and b^2.
// T: storage type of vol and out (float, or uint16_t for BF16); b and invRow are FP32
// and the ray sum is accumulated in FP32.
// ---------------------------------------------------------------------------
// NJ in-plane cells x 2 or 3 z taps of a slice footprint, all loads issued before the
// first is used: the cells' weights times their mask bytes already taken (wy0, wy1, wy2)
// and the z taps read as the fixed run o, o + 1, o + 2 of cell ij (o = ob + ij * cS + k0;
// one offset per cell, the steps in the loads' immediate offsets, all range-checked: a tap
// of zero weight reads a neighbouring element of the array, finite, or past its end 0, and
// adds an exact zero).  NT = 2 when no lane needs a third z cell (the term it drops is an
// exact zero).  The combining order of the general loop in fp_body.
// BF16 volumes read each run as the 8 bytes from the even element kb = k0 & ~1 on (one
// aligned 4-element load; ob even), the three taps picked by a funnel shift by 16 (k0 - kb)
// bits.
template <int NJ, int NT, typename TV>
__device__ __forceinline__ float fp_cells_w(const Buf<TV>& V, int ob, int cS, int k0,
                                            float wz0, float wz1, float wz2, float wy0,
                                            float wy1, float wy2) {
  float a[NJ][3];
  if constexpr (sizeof(TV) == 2) {
    const int kb = k0 & ~1;
    const uint32_t sh = (uint32_t)(k0 & 1) << 4;
    uint2 w[NJ];
#pragma unroll
    for (int ij = 0; ij < NJ; ++ij) w[ij] = V.load8b(kb + (ob + ij * cS));
#pragma unroll
    for (int ij = 0; ij < NJ; ++ij) {
      // elements k0, k0 + 1 as the low / high halves of x, k0 + 2 as the low half of y
      const uint32_t x = __builtin_amdgcn_alignbit(w[ij].y, w[ij].x, sh);
      a[ij][0] = __uint_as_float(x << 16);
      a[ij][1] = __uint_as_float(x & 0xffff0000u);
      if constexpr (NT == 3) a[ij][2] = __uint_as_float((w[ij].y >> sh) << 16);
    }
  } else {
#pragma unroll
    for (int ij = 0; ij < NJ; ++ij) {
      const int o = k0 + (ob + ij * cS);
      a[ij][0] = V[o];
      a[ij][1] = V[o + 1];
      if constexpr (NT == 3) a[ij][2] = V[o + 2];
    }
  }
  float s = 0.f;
#pragma unroll
  for (int ij = 0; ij < NJ; ++ij) {
    const float wy = ij == 0 ? wy0 : (ij == 1 ? wy1 : wy2);
    float g = wz0 * a[ij][0];
    g = fmaf(wz1, a[ij][1], g);
    if constexpr (NT == 3) g = fmaf(wz2, a[ij][2], g);
    s = fmaf(wy, g, s);
  }
  return s;
}

// In-plane part of slice ii of one detector column's footprint: the boundary lines
// t_b(ii) = a_b + ii f_b of its two boundaries (FP64: positions run to the slice width,
// where FP32 keeps only ~3e-5 of a cell), the footprint [ylo, yhi] clipped to the mask's
// extent e on the slice, its cells j0 .. j0 + nj - 1 with first / last overlap weights
// (FP64 differences, rounded once), the z magnification r and 1 / ((yhi - ylo) r).
// False when the clipped footprint is empty.
struct SliceIP {
  float r, inv, wj0, wjl;
  int j0, nj;
};
__device__ __forceinline__ bool fp_inplane(const double2& bl, const double2& br, const float4& ct,
                                           const float2& e, int ii, int NI, SliceIP& o) {
  const double fi = (double)ii;
  const double tl = fma(fi, bl.y, bl.x), tr = fma(fi, br.y, br.x);
  const double ylo = fmin(tl, tr), yhi = fmax(tl, tr);
  const double clo = fmax(ylo, (double)e.x), chi = fmin(yhi, (double)e.y);
  if (!(clo < chi)) return false;
  o.r = fmaf((float)ii, ct.y, ct.x);
  o.inv = __builtin_amdgcn_rcpf((float)(yhi - ylo) * o.r);
  o.j0 = (int)floor(clo + 0.5);
  o.nj = min((int)floor(chi + 0.5), NI - 1) - o.j0 + 1;
  o.wj0 = (float)(o.nj == 1 ? chi - clo : ((double)o.j0 + 0.5) - clo);
  o.wjl = (float)(chi - ((double)(o.j0 + o.nj - 1) - 0.5));
  return true;
}

// ONES: the projection of a volume of ones (the row sums of OS-SART), read from no
// volume: the z weights of a footprint sum to its clipped height zh - zl, so a slice
// adds (zh - zl) * sum_j wy_j mask_j (the same sum in another rounding order).
template <bool FUSED, bool COLW, typename TV, typename TO, bool ONES = false>
__device__ __forceinline__ void fp_body(const FpArgs<TV, TO>& A, int k, int ublk, double* slot) {
  const Dims& d = A.d;
  const int p = A.vb + k * A.vs;
  const int nray = d.DNU * d.DNV;
  // COLW (DNV == 64): a wave is exactly one detector column, so everything in-plane
  // is uniform and lives in scalar registers
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ray = COLW ? (ublk * kWaves + wid) * 64 + (threadIdx.x & 63) : ublk * kBlock + threadIdx.x;
  double res = 0.0, bsq = 0.0;
  if (ray < nray) {
    const int u = COLW ? (ublk * kWaves + wid) : ray / d.DNV;
    const int v = ray - u * d.DNV;
    const ViewP vp = gload(A.views, p);
    const double2 bl = gload(A.bnd64, (size_t)p * (d.DNU + 1) + u);
    const double2 br = gload(A.bnd64, (size_t)p * (d.DNU + 1) + u + 1);
    const float4 ct = gload(A.cols, (size_t)p * d.DNU + u);
    const float Bv = gp(A.Bz)[v], Bv1 = gp(A.Bz)[v + 1];
    int NS, NI, sS, cS, mS, mC;
    if (vp.xmarch) {
      NS = d.XN; NI = d.YN; sS = d.ZN; cS = d.XN * d.ZN; mS = 1; mC = d.XN;
    } else {
      NS = d.YN; NI = d.XN; sS = d.XN * d.ZN; cS = d.ZN; mS = d.XN; mC = 1;
    }
    // z in cell-index units relative to azi (exact shifts of the volume's bounds)
    const float zlo_b = -0.5f - (float)vp.azi, zhi_b = ((float)d.ZN - 0.5f) - (float)vp.azi;
    const int ZN1 = d.ZN - 1;
    const float2* me = A.mext + (vp.xmarch ? 0 : d.XN);
    const int NV = d.XN * d.YN * d.ZN;
    const Buf<TV> V(A.vol, (uint32_t)NV);
    // BF16 runs as aligned 8-byte loads: even column starts on a 4-byte aligned array
    const bool bf16_runs = d.ZN % 2 == 0 && (reinterpret_cast<uintptr_t>(A.vol) & 3) == 0;
    const CTR_GLOBAL float* maskf = gp(A.maskf);
    float acc = 0.f;
    // one slice: footprint [clo, chi] x [zl, zh] over the cells j0 .. j0 + nj - 1
    auto slice = [&](int ii, float r, float inv, int j0, int nj, float wj0, float wjl) {
      float zl = fmaf(r, Bv, vp.azf), zh = fmaf(r, Bv1, vp.azf);
      zl = fmaxf(zl, zlo_b);
      zh = fminf(zh, zhi_b);
      const bool zok = zl < zh;
      if (!__any(zok)) return;
      if constexpr (ONES) {
        float W = 0.f;
        for (int ij = 0; ij < nj; ++ij)
          W = fmaf((ij == 0) ? wj0 : (ij == nj - 1 ? wjl : 1.f), maskf[ii * mS + (j0 + ij) * mC],
                   W);
        if (zok) acc = fmaf((zh - zl) * W, inv, acc);
        return;
      }
      // first cell k0 = floor(zl + .5) and its upper edge e; with dk = zh - e the overlaps
      // of cells k0, k0 + 1, k0 + 2 are min(zh, e) - zl, clamp(dk, 0, 1), clamp(dk - 1, 0, 1)
      const float fk = floorf(zl + 0.5f);
      const float ek = fk + 0.5f;
      const int k0 = min((int)fk + vp.azi, ZN1);
      const float dk = zok ? zh - ek : -1.f;
      const float wz0 = zok ? fminf(zh, ek) - zl : 0.f;
      const float wz1 = __builtin_amdgcn_fmed3f(dk, 0.f, 1.f);
      const float wz2 = __builtin_amdgcn_fmed3f(dk - 1.f, 0.f, 1.f);
      int o = ii * sS + j0 * cS + k0;
      int mo = ii * mS + j0 * mC;
      // taps of zero weight re-read the first cell, never a neighbouring column
      const int t1 = dk > 0.f ? 1 : 0, t2 = dk > 1.f ? 2 : 0;
      float s = 0.f;
      if (!__any(dk > 2.f)) {
        for (int ij = 0; ij < nj; ++ij) {
          const float wy = ((ij == 0) ? wj0 : (ij == nj - 1 ? wjl : 1.f)) * maskf[mo];
          float g = wz0 * V[o];
          g = fmaf(wz1, V[o + t1], g);
          g = fmaf(wz2, V[o + t2], g);
          s = fmaf(wy, g, s);
          o += cS;
          mo += mC;
        }
      } else {
        const int nk = zok ? min((int)floorf(zh + 0.5f) + vp.azi, ZN1) - k0 + 1 : 0;
        float wf, wzl;
        edge_weights(zl, zh, k0 - vp.azi, nk, wf, wzl);
        for (int ij = 0; ij < nj; ++ij) {
          const float wy = ((ij == 0) ? wj0 : (ij == nj - 1 ? wjl : 1.f)) * maskf[mo];
          float g = 0.f;
          for (int ik = 0; ik < nk; ++ik)
            g = fmaf((ik == 0) ? wz0 : (ik == nk - 1 ? wzl : 1.f), V[o + ik], g);
          s = fmaf(wy, g, s);
          o += cS;
          mo += mC;
        }
      }
      acc = fmaf(s, inv, acc);
    };
    if constexpr (COLW && !ONES) {
      // the in-plane part of a slice is uniform over the wave: lane l evaluates it for
      // slice ii0 + l (footprint, 1 / its area, and each cell's edge weight times its mask
      // byte) into the wave's LDS table, and the loop over the slices the footprint meets
      // reads a slice's entry at a uniform address (same arithmetic per slice)
      __shared__ float4 s_slices[kWaves * 128];
      float4* rec = s_slices + wid * 128;
      const int lane = threadIdx.x & 63;
      for (int ii0 = 0; ii0 < NS; ii0 += 64) {
        const int il = ii0 + lane;
        float q_r = 0.f, q_inv = 0.f, q_wj0 = 0.f, q_wjl = 0.f, wy0 = 0.f, wy1 = 0.f, wy2 = 0.f;
        int q_j0 = 0, q_nj = 0;
        bool ok = false;
        SliceIP si;
        // clipped to the mask's extent on the slice (inside [-0.5, NI - 0.5]): cells
        // outside it are zero
        if (il < NS && fp_inplane(bl, br, ct, gload(me, il), il, NI, si)) {
          ok = true;
          {
            q_r = si.r;
            q_inv = si.inv;
            q_j0 = si.j0;
            q_nj = si.nj;
            q_wj0 = si.wj0;
            q_wjl = si.wjl;
            const int mo = il * mS + q_j0 * mC;
            wy0 = q_wj0 * maskf[mo];
            if (q_nj >= 2) wy1 = (q_nj == 2 ? q_wjl : 1.f) * maskf[mo + mC];
            if (q_nj >= 3) wy2 = q_wjl * maskf[mo + 2 * mC];
          }
        }
        rec[2 * lane] = make_float4(q_r, q_inv, wy0, wy1);
        rec[2 * lane + 1] = make_float4(wy2, q_wj0, q_wjl, 0.f);
        uint64_t live = __ballot(ok);
        while (live) {
          const int j = __builtin_ctzll(live);
          live &= live - 1;
          const int ii = ii0 + j;
          const int j0 = __builtin_amdgcn_readlane(q_j0, j);
          const int nj = __builtin_amdgcn_readlane(q_nj, j);
          const float4 ra = rec[2 * j], rb = rec[2 * j + 1];
          const float r = ra.x;
          float zl = fmaf(r, Bv, vp.azf), zh = fmaf(r, Bv1, vp.azf);
          zl = fmaxf(zl, zlo_b);
          zh = __builtin_elementwise_minimum(zh, zhi_b);
          const bool zok = zl < zh;
          if (!__any(zok)) continue;
          // first cell k0 = floor(zl + .5) (+ azi) and its upper edge e; with dk = zh - e
          // the overlaps of cells k0, k0 + 1, k0 + 2 are min(zh, e) - zl, clamp(dk, 0, 1),
          // clamp(dk - 1, 0, 1)
          const float fk = floorf(zl + 0.5f);
          const float ek = fk + 0.5f;
          const int k0 = min((int)fk + vp.azi, ZN1);
          // outside [-0.5, ZN - 0.5] (zl >= zh): dk < 0 and wz0 <= 0 clamped, so every
          // weight is zero
          const float dk = zh - ek;
          const float wz0 = fmaxf(__builtin_elementwise_minimum(zh, ek) - zl, 0.f);
          const float wz1 = __builtin_amdgcn_fmed3f(dk, 0.f, 1.f);
          const float wz2 = __builtin_amdgcn_fmed3f(dk - 1.f, 0.f, 1.f);
          const int ob = ii * sS + j0 * cS;  // wave-uniform
          float s = 0.f;
          if (!__any(dk > 2.f)) {
            // the fixed runs of fp_cells_w stay inside the array (its merged multi-element
            // loads never straddle the end): not on the array's last column
            if (nj <= 3 && ob + (nj - 1) * cS + d.ZN + 3 <= NV &&
                (sizeof(TV) == 4 || bf16_runs)) {
              // nj is uniform: straight-line code per cell count, all loads first
              if (!__any(dk > 1.f)) {  // at most two z cells in every lane's footprint
                switch (nj) {
                  case 1: s = fp_cells_w<1, 2>(V, ob, cS, k0, wz0, wz1, wz2, ra.z, ra.w, rb.x); break;
                  case 2: s = fp_cells_w<2, 2>(V, ob, cS, k0, wz0, wz1, wz2, ra.z, ra.w, rb.x); break;
                  default: s = fp_cells_w<3, 2>(V, ob, cS, k0, wz0, wz1, wz2, ra.z, ra.w, rb.x); break;
                }
              } else {
                switch (nj) {
                  case 1: s = fp_cells_w<1, 3>(V, ob, cS, k0, wz0, wz1, wz2, ra.z, ra.w, rb.x); break;
                  case 2: s = fp_cells_w<2, 3>(V, ob, cS, k0, wz0, wz1, wz2, ra.z, ra.w, rb.x); break;
                  default: s = fp_cells_w<3, 3>(V, ob, cS, k0, wz0, wz1, wz2, ra.z, ra.w, rb.x); break;
                }
              }
            } else {
              // taps of zero weight re-read the first cell
              const int t1 = dk > 0.f ? 1 : 0, t2 = dk > 1.f ? 2 : 0;
              int mo = ii * mS + j0 * mC;
              int o = ob + k0;
              for (int ij = 0; ij < nj; ++ij) {
                const float wy = ((ij == 0) ? rb.y : (ij == nj - 1 ? rb.z : 1.f)) * maskf[mo];
                float g = wz0 * V[o];
                g = fmaf(wz1, V[o + t1], g);
                g = fmaf(wz2, V[o + t2], g);
                s = fmaf(wy, g, s);
                o += cS;
                mo += mC;
              }
            }
          } else {
            const int nk = zok ? min((int)floorf(zh + 0.5f) + vp.azi, ZN1) - k0 + 1 : 0;
            float wf, wzl;
            edge_weights(zl, zh, k0 - vp.azi, nk, wf, wzl);
            int mo = ii * mS + j0 * mC;
            int o = ob + k0;
            for (int ij = 0; ij < nj; ++ij) {
              const float wy = ((ij == 0) ? rb.y : (ij == nj - 1 ? rb.z : 1.f)) * maskf[mo];
              float g = 0.f;
              for (int ik = 0; ik < nk; ++ik)
                g = fmaf((ik == 0) ? wz0 : (ik == nk - 1 ? wzl : 1.f), V[o + ik], g);
              s = fmaf(wy, g, s);
              o += cS;
              mo += mC;
            }
          }
          acc = fmaf(s, ra.y, acc);
        }
      }
    } else if constexpr (COLW) {
      // ONES: lane l evaluates the in-plane part of slice ii0 + l, with the sum W of its
      // cells' edge weights times mask bytes (the order of slice()), read back with
      // readlane; a slice adds (zh - zl) W / area as slice() does
      const int lane = threadIdx.x & 63;
      for (int ii0 = 0; ii0 < NS; ii0 += 64) {
        const int il = ii0 + lane;
        float q_r = 0.f, q_inv = 0.f, q_W = 0.f;
        bool ok = false;
        SliceIP si;
        if (il < NS && fp_inplane(bl, br, ct, gload(me, il), il, NI, si)) {
          ok = true;
          q_r = si.r;
          q_inv = si.inv;
          for (int ij = 0; ij < si.nj; ++ij)
            q_W = fmaf((ij == 0) ? si.wj0 : (ij == si.nj - 1 ? si.wjl : 1.f),
                       maskf[il * mS + (si.j0 + ij) * mC], q_W);
        }
        uint64_t live = __ballot(ok);
        while (live) {
          const int j = __builtin_ctzll(live);
          live &= live - 1;
          const float r = rdl(q_r, j);
          float zl = fmaf(r, Bv, vp.azf), zh = fmaf(r, Bv1, vp.azf);
          zl = fmaxf(zl, zlo_b);
          zh = fminf(zh, zhi_b);
          const bool zok = zl < zh;
          if (!__any(zok)) continue;
          const float W = rdl(q_W, j), inv = rdl(q_inv, j);
          if (zok) acc = fmaf((zh - zl) * W, inv, acc);
        }
      }
    } else {
      for (int ii = 0; ii < NS; ++ii) {
        SliceIP si;
        if (fp_inplane(bl, br, ct, gload(me, ii), ii, NI, si))
          slice(ii, si.r, si.inv, si.j0, si.nj, si.wj0, si.wjl);
      }
    }
    const float K = ct.w * __builtin_amdgcn_sqrtf(ct.z + gp(A.h2)[v]) * gp(A.invdB)[v];
    const float val = acc * K;
    const size_t o = ((size_t)k * d.DNU + u) * d.DNV + v;
    if constexpr (FUSED) {
      const size_t og = ((size_t)p * d.DNU + u) * d.DNV + v;
      const float bb = gp(A.b)[og];
      const float diff = bb - val;
      if (A.out) st(gp(A.out), o, A.invRow ? diff * gp(A.invRow)[o] : diff);
      res = (double)diff * (double)diff;
      bsq = (double)bb * (double)bb;
    } else {
      st(gp(A.out), o, val);
    }
  }
  if constexpr (FUSED) block_sum2(res, bsq, slot);
}

template <bool FUSED, bool COLW, typename TV, typename TO>
__global__ __launch_bounds__(kBlock) void fp_kernel(FpArgs<TV, TO> A) {
  const int hw = blockIdx.y * gridDim.x + blockIdx.x;
  const int lb = xcd_logical_block(hw, gridDim.x * gridDim.y);
  const int k = lb / gridDim.x;
  fp_body<FUSED, COLW, TV, TO>(A, k, lb - k * gridDim.x, A.partials + 2 * hw);
}

// Forward projection of a volume of ones (A.vol unused).
template <bool COLW>
__global__ __launch_bounds__(kBlock) void fp_ones_kernel(FpArgs<float, float> A) {
  const int hw = blockIdx.y * gridDim.x + blockIdx.x;
  const int lb = xcd_logical_block(hw, gridDim.x * gridDim.y);
  const int k = lb / gridDim.x;
  fp_body<false, COLW, float, float, true>(A, k, lb - k * gridDim.x, nullptr);
}

// Batched persistent forward projection: each workgroup takes one contiguous range of
// the flat work list (XCD-aware, as in fp_kernel), items of one scan being
// consecutive (view-major, then column blocks).
// Work assignment of the batched kernels (sched): the workgroup's first item, the item
// step and the end.
//  0 contiguous: one range per workgroup, ranges ordered per XCD (neighbouring items
//    share an L2);
//  1 strided: items lb, lb + G, ... (every workgroup samples the whole list, which
//    balances uneven item costs);
//  2 one workgroup per item (G == total): workgroups are dispatched in order, so the
//    items of one scan run together on all XCDs, and within a scan's range the
//    workgroup -> item map is the XCD-aware map of the single-scan kernels.
template <typename Item>
__device__ __forceinline__ void batch_range(const Item* __restrict__ D, int nscan, int total,
                                            int sched, int& i0, int& step, int& i1) {
  if (sched == 2) {
    const int hw = blockIdx.x;
    int s = 0;
    while (s + 1 < nscan && D[s + 1].begin <= hw) ++s;
    const int b0 = D[s].begin;
    const int n = (s + 1 < nscan ? D[s + 1].begin : total) - b0;
    i0 = b0 + xcd_logical_block(hw - b0, n);
    step = 1;
    i1 = i0 + 1;
    return;
  }
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  if (sched == 1) {
    i0 = lb;
    step = gridDim.x;
    i1 = total;
  } else {
    const int per = (total + gridDim.x - 1) / gridDim.x;
    i0 = lb * per;
    step = 1;
    i1 = min(total, i0 + per);
  }
}

template <bool FUSED, bool COLW, typename T>
__global__ __launch_bounds__(kBlock) void fp_batch_kernel(const FpBatchItem<T>* __restrict__ D,
                                                          int nscan, int total, int sched) {
  int i0, step, i1;
  batch_range(D, nscan, total, sched, i0, step, i1);
  int s = 0;
  for (int item = i0; item < i1; item += step) {
    if (sched == 1) s = 0;
    while (s + 1 < nscan && D[s + 1].begin <= item) ++s;
    const FpBatchItem<T>& B = D[s];
    const int local = item - B.begin;
    const int k = local / B.upv;
    fp_body<FUSED, COLW, T, T>(B.a, k, local - k * B.upv, B.a.partials + 2 * local);
  }
}


// fp_batch_kernel for the one-workgroup-per-item schedule (grid = total): no item loop.
template <bool FUSED, bool COLW, typename T>
__global__ __launch_bounds__(kBlock) void fp_batch1_kernel(const FpBatchItem<T>* __restrict__ D,
                                                           int nscan, int total) {
  int i0, step, i1;
  batch_range(D, nscan, total, 2, i0, step, i1);
  int s = 0;
  while (s + 1 < nscan && D[s + 1].begin <= i0) ++s;
  const FpBatchItem<T>& B = D[s];
  const int local = i0 - B.begin;
  const int k = local / B.upv;
  fp_body<FUSED, COLW, T, T>(B.a, k, local - k * B.upv, B.a.partials + 2 * local);
}

// ---------------------------------------------------------------------------
// Reference back projector: one wave per (x, y) column, each lane owning ZG
// voxels z = z0 + lane + 64 g (g < ZG), views looped inside.
// Reference: DD3_gpu_back_ker<_BRANCHLESS>, iterativeRecon/DD3_GPU.cu:1397-1475.
// The voxel's shadow in (fan-angle index, row index) space is integrated over the
// piecewise-constant sinogram and weighted by cosVal.  The fan-angle coordinate is
// the same for the whole column and is resolved in FP64 once per (column, view);
// the row coordinates are per lane in FP32.  A view whose row shadow misses the
// detector for every lane of a group is skipped for that group.
// UPDATE: vol += acc * invCol (os_sart.m:32,37-40), and shadow (when not NULL) gets
// the updated value rounded to BF16; otherwise vol = acc.
// T: storage type of prj (float, or uint16_t for BF16); vol is FP32.
// ---------------------------------------------------------------------------
// Per-(column, view) constants of the reference back projector, uniform over a wave:
// the fan-angle shadow (FP64) as first cell u0, cell count nu and first / last column
// weights, the cell-width factor with the sign of the corner difference folded in, and
// the row-shadow constants (rows of a voxel at dZ = Pz - sz: [dZ rlS + cA, dZ rlS + cB]).
struct BpView {
  float rlS, cA, cB, sz, l2, cws, wu0, fh;
  int u0, nu;
  bool uvis;
};

template <typename T>
__device__ __forceinline__ BpView bp_view(const BpArgs<T>& A, int p, double Px, double Py,
                                          float Pxf, float Pyf) {
  const Dims& d = A.d;
  const BpConsts& c = A.c;
  const BpConsts64& c64 = A.c64;
  BpView r{};
  const ViewP vp = gload(A.views, p);
  // row shadow, FP32 (DD3_GPU.cu:1446-1450): t = (dZ +- hz) * S / l + cV
  const float dXf = Pxf - vp.sx, dYf = Pyf - vp.sy;
  const float l2f = dXf * dXf + dYf * dYf;
  const float rlS = __builtin_amdgcn_rsqf(l2f) * c.S;
  const bool rneg = rlS < 0.f;  // rows then run backwards: swap the bounds
  // fan-angle shadow, FP64 (DD3_GPU.cu:1447-1464); sy*dX - sx*dY == sy*Px - sx*Py
  const double2 s64 = gload(A.src64, p);
  const double dX = Px - s64.x, dY = Py - s64.y;
  const double l2 = dX * dX + dY * dY;
  const double rl = rsqrt(l2);
  const double alpha = asin_fan((s64.y * Px - s64.x * Py) * c64.inv_sid * rl);
  const double w = vp.wx ? dX : dY;
  const double dA = w * (rl * rl) * (c64.dx * 0.5);
  const double tL = (alpha - dA) * c64.ib + c64.cU;
  const double tR = (alpha + dA) * c64.ib + c64.cU;
  const double ulo = fmax(fmin(tL, tR), -0.5);
  const double uhi = fmin(fmax(tL, tR), (double)d.DNU - 0.5);
  r.uvis = ulo < uhi;
  if (r.uvis) {
    const int u0 = (int)floor(ulo + 0.5);
    const int nu = min((int)floor(uhi + 0.5), d.DNU - 1) - u0 + 1;
    // column weights: first / last partial, interior 1 (fractions taken in FP64)
    const float fl = (float)(((double)u0 + 0.5) - ulo);
    r.fh = (float)(uhi - ((double)(u0 + nu - 1) - 0.5));
    r.wu0 = (nu == 1) ? (float)(uhi - ulo) : fl;
    // sign of the reference's corner difference, (tR - tL) and (tU - tD), folded
    // into the cell-width factor (a sign flip is exact)
    const float sgn = ((tR < tL) != rneg) ? -1.f : 1.f;
    r.cws = sgn * (c.dx * __builtin_amdgcn_rcpf((float)w));
    r.u0 = u0;
    r.nu = nu;
  }
  const float cUp = fmaf(c.hz, rlS, c.cV), cDp = fmaf(-c.hz, rlS, c.cV);
  r.cA = rneg ? cUp : cDp;  // lower row bound
  r.cB = rneg ? cDp : cUp;  // upper row bound
  r.rlS = rlS;
  r.sz = vp.sz;
  r.l2 = l2f;
  return r;
}

// NU columns x 3 row taps of a shadow: all 3 NU loads are issued before the first is
// used (NU is the uniform column count, so every case is straight-line code), then the
// taps are combined in the order of the general loop below.
// NT = 2 when no lane of the wave needs a third row (wv2 == 0 everywhere: the term it
// drops is an exact zero).
template <int NU, int NT, typename T>
__device__ __forceinline__ float bp_cells(const Buf<T>& pp, int o, int DNV, int t1, int t2,
                                          float wv0, float wv1, float wv2, float wu0, float fh) {
  // no contraction: an interior column's weight 1 would otherwise fuse its add with the
  // first column's product, which the staged rows of bp_win_body (a run-time weight)
  // round separately
#pragma clang fp contract(off)
  float a[NU][3];
#pragma unroll
  for (int iu = 0; iu < NU; ++iu) {
    if constexpr (NT == 3) {
      pp.tap3(o + iu * DNV, t1, t2, a[iu][0], a[iu][1], a[iu][2]);
    } else {
      a[iu][0] = pp[o + iu * DNV];
      a[iu][1] = pp[o + iu * DNV + t1];
    }
  }
  // rows first combined over the columns (R_t = sum_u wu y[u][v0 + t]), then over the rows
  float R[3];
#pragma unroll
  for (int t = 0; t < NT; ++t) R[t] = wu0 * a[0][t];
#pragma unroll
  for (int iu = 1; iu < NU; ++iu) {
    const float wu = (iu == NU - 1) ? fh : 1.f;
#pragma unroll
    for (int t = 0; t < NT; ++t) R[t] = fmaf(wu, a[iu][t], R[t]);
  }
  float s = wv0 * R[0];
  s = fmaf(wv1, R[1], s);
  if constexpr (NT == 3) s = fmaf(wv2, R[2], s);
  return s;
}

// The voxel's shadow [vlo, vhi] (rows) x cells u0 .. u0 + nu - 1 integrated over the
// piecewise-constant view pp (first / last column weights wu0, fh).
template <typename T>
__device__ __forceinline__ float bp_cell_sum(const Buf<T>& pp, int u0, int nu, float wu0,
                                             float fh, float vlo, float vhi, int DNV) {
#pragma clang fp contract(off)
  // first row v0 = floor(vlo + .5) and its upper edge e; with d = vhi - e the overlaps of
  // rows v0, v0 + 1, v0 + 2 are min(vhi, e) - vlo, clamp(d, 0, 1), clamp(d - 1, 0, 1)
  // (the first / interior / last weights of edge_weights, without the row count)
  const float fv = floorf(vlo + 0.5f);
  const float e = fv + 0.5f;
  const int v0 = (int)fv;
  const float d = vhi - e;
  float s = 0.f;
  if (nu <= 4 && d <= 2.f) {
    const float wv0 = fminf(vhi, e) - vlo;
    const float wv1 = __builtin_amdgcn_fmed3f(d, 0.f, 1.f);
    const float wv2 = __builtin_amdgcn_fmed3f(d - 1.f, 0.f, 1.f);
    // taps of zero weight re-read the first row (never a row outside the shadow)
    const int t1 = d > 0.f ? 1 : 0, t2 = d > 1.f ? 2 : 0;
    const int o = u0 * DNV + v0;
    if (!__any(d > 1.f)) {  // at most two rows in every lane's shadow
      switch (nu) {  // uniform over the wave
        case 1: s = bp_cells<1, 2>(pp, o, DNV, t1, t2, wv0, wv1, wv2, wu0, fh); break;
        case 2: s = bp_cells<2, 2>(pp, o, DNV, t1, t2, wv0, wv1, wv2, wu0, fh); break;
        case 3: s = bp_cells<3, 2>(pp, o, DNV, t1, t2, wv0, wv1, wv2, wu0, fh); break;
        default: s = bp_cells<4, 2>(pp, o, DNV, t1, t2, wv0, wv1, wv2, wu0, fh); break;
      }
    } else {
      switch (nu) {
        case 1: s = bp_cells<1, 3>(pp, o, DNV, t1, t2, wv0, wv1, wv2, wu0, fh); break;
        case 2: s = bp_cells<2, 3>(pp, o, DNV, t1, t2, wv0, wv1, wv2, wu0, fh); break;
        case 3: s = bp_cells<3, 3>(pp, o, DNV, t1, t2, wv0, wv1, wv2, wu0, fh); break;
        default: s = bp_cells<4, 3>(pp, o, DNV, t1, t2, wv0, wv1, wv2, wu0, fh); break;
      }
    }
  } else {
    const int nv = min((int)floorf(vhi + 0.5f), DNV - 1) - v0 + 1;
    float wv0, wvl;
    edge_weights(vlo, vhi, v0, nv, wv0, wvl);
    for (int iv = 0; iv < nv; ++iv) {
      const float wv = (iv == 0) ? wv0 : (iv == nv - 1 ? wvl : 1.f);
      const int o = u0 * DNV + v0 + iv;
      float R = wu0 * pp[o];
      for (int iu = 1; iu < nu; ++iu) R = fmaf((iu == nu - 1) ? fh : 1.f, pp[o + iu * DNV], R);
      s = fmaf(wv, R, s);
    }
  }
  return s;
}

// Staged windowed back projector (DNV <= kStageRows): per wave, the current view's
// combined rows (kStagePad floats: two rows of zeros past the last row for taps there),
// then a table of 64 view records (kRecF floats each: row-shadow constants, cosVal
// constants and the four column weights, written by lane l for view k0 + l).  kStageF
// floats per wave, a multiple of 4 so every wave's table is 16-byte aligned.
constexpr int kStageRows = 64;
constexpr int kStagePad = kStageRows + 2;
constexpr int kRowsF = kStagePad + 2;  // rows, then the record table 16-byte aligned
constexpr int kRecF = 12;
constexpr int kStageF = kRowsF + 64 * kRecF;
// staged waves' column sums carry 64 floats past the column: a pass's lanes past the
// column's end (whose shadows need not be empty) add there, and the passes need no
// per-voxel bound test
constexpr int kLaccPad = 64;
static_assert(kRowsF % 4 == 0 && kStageF % 4 == 0, "per-wave stage: 16-byte multiples");

// bp_cell_sum over a view whose columns are already combined per row in LDS
// (lR[v] = sum_u wu y[u][v], rows past DNV zero, two zero rows past the stage), with the
// same sum: the row weights min(vhi, e) - vlo, clamp(d, 0, 1), clamp(d - 1, 0, 1) of a
// shadow over at most three rows are exactly the first / interior / last weights of
// edge_weights (d = vhi - e and d - 1 are exact differences), and taps past the shadow
// add exact zeros (the third row's is added in every pass: most passes have a lane that
// needs it), so the choice between the two forms may be made for the whole wave.
__device__ __forceinline__ float bp_lds_sum(const float* lR, float vlo, float vhi, int DNV) {
#pragma clang fp contract(off)
  const float fv = floorf(vlo + 0.5f);
  const float e = fv + 0.5f;
  const int v0 = (int)fv;
  const float d = vhi - e;
  float s;
  if (!__any(d > 2.f)) {
    const float wv0 = __builtin_elementwise_minimum(vhi, e) - vlo;
    const float wv1 = __builtin_amdgcn_fmed3f(d, 0.f, 1.f);
    s = wv0 * lR[v0];
    s = fmaf(wv1, lR[v0 + 1], s);
    s = fmaf(__builtin_amdgcn_fmed3f(d - 1.f, 0.f, 1.f), lR[v0 + 2], s);
  } else {
    const int nv = min((int)floorf(vhi + 0.5f), DNV - 1) - v0 + 1;
    float wv0, wvl;
    edge_weights(vlo, vhi, v0, nv, wv0, wvl);
    s = 0.f;
    for (int iv = 0; iv < nv; ++iv)
      s = fmaf((iv == 0) ? wv0 : (iv == nv - 1 ? wvl : 1.f), lR[v0 + iv], s);
  }
  return s;
}

// Final value of voxel z of a column: vol = acc, or the OS-SART update
// vol += acc * invCol (and its BF16 shadow).
template <bool UPDATE, typename T>
__device__ __forceinline__ void bp_store(const BpArgs<T>& A, CTR_GLOBAL float* vcol, int col,
                                         int z, float acc) {
  if constexpr (UPDATE) {
    const size_t i = (size_t)col * A.d.ZN + z;
    const float nx = fmaf(acc, gp(A.invCol)[i], vcol[z]);
    vcol[z] = nx;
    if (A.shadow) gp(A.shadow)[i] = to_bf16_dither(nx, i);
  } else {
    vcol[z] = acc;
  }
}

template <bool UPDATE, int ZG, typename T>
__device__ __forceinline__ void bp_ref_body(const BpArgs<T>& A, int col, int zblk) {
  const Dims& d = A.d;
  const BpConsts64& c64 = A.c64;
  const int lane = threadIdx.x & 63;
  if (col >= d.XN * d.YN) return;
  const int zb0 = zblk * (64 * ZG);  // first z of the wave (uniform)
  const int zb = zb0 + lane;
  const int y = col / d.XN, x = col - y * d.XN;
  CTR_GLOBAL float* vcol = gp(A.vol) + (size_t)col * d.ZN;
  if (gp(A.mask)[col] != 1) {
    if constexpr (!UPDATE) {
#pragma unroll
      for (int g = 0; g < ZG; ++g)
        if (zb + 64 * g < d.ZN) vcol[zb + 64 * g] = 0.f;
    }
    return;
  }
  const double Px = ((double)x - c64.ocx) * c64.dx;
  const double Py = ((double)y - c64.ocy) * c64.dx;
  const float Pxf = (float)Px, Pyf = (float)Py;
  // z extent of the wave's voxel centres (uniform), for the whole-column view test
  const float Zlo = ((float)zb0 - d.ocz) * d.dz;
  const float Zhi = ((float)min(zb0 + 64 * ZG, d.ZN) - 1.f - d.ocz) * d.dz;
  float acc[ZG];
#pragma unroll
  for (int g = 0; g < ZG; ++g) acc[g] = 0.f;
  const float DNVh = (float)d.DNV - 0.5f;
  const int DNV = d.DNV;
  const uint32_t frame = (uint32_t)(d.DNU * d.DNV);
  // Views are taken 64 at a time: lane l first evaluates everything of view k0 + l that
  // is uniform over the column (visibility, fan-angle shadow in FP64, column weights,
  // row-shadow constants); the loop over the visible views then reads them with
  // readlane into scalar registers.  The arithmetic per view is unchanged.
  for (int k0 = 0; k0 < A.vc; k0 += 64) {
    const int kl = k0 + lane;
    BpView q{};
    bool vis = false;
    if (kl < A.vc) {
      q = bp_view(A, A.vb + kl * A.vs, Px, Py, Pxf, Pyf);
      // the whole group range of rows, to skip views the wave's voxels do not meet
      const float ta = fmaf(Zlo - q.sz, q.rlS, q.cA), tb = fmaf(Zhi - q.sz, q.rlS, q.cB);
      vis = q.uvis && fmaxf(ta, tb) > -0.5f && fminf(ta, tb) < DNVh;
    }
    uint64_t live = __ballot(vis);
    while (live) {
      const int j = __builtin_ctzll(live);
      live &= live - 1;
      const int k = k0 + j;
      const float rlS = rdl(q.rlS, j), cA = rdl(q.cA, j), cB = rdl(q.cB, j);
      const float vsz = rdl(q.sz, j), l2f = rdl(q.l2, j), cws = rdl(q.cws, j);
      const float wu0 = rdl(q.wu0, j), fh = rdl(q.fh, j);
      const int u0 = __builtin_amdgcn_readlane(q.u0, j), nu = __builtin_amdgcn_readlane(q.nu, j);
      const Buf<T> pp(A.prj + (size_t)k * frame, frame);
#pragma unroll
      for (int g = 0; g < ZG; ++g) {
        // voxel centre offset as in bp_win_body: one explicit fma, so that neither body
        // can round it differently (a hoisted product would be rounded on its own)
        const float dZ = fmaf((float)(zb + 64 * g) - d.ocz, d.dz, -vsz);
        const float vlo = fmaxf(fmaf(dZ, rlS, cA), -0.5f);
        const float vhi = fminf(fmaf(dZ, rlS, cB), DNVh);
        const bool vok = zb + 64 * g < d.ZN && vlo < vhi;
        if (!__any(vok)) continue;
        if (!vok) continue;
        const float cosVal = cws * __builtin_amdgcn_sqrtf(fmaf(dZ, dZ, l2f));
        acc[g] = fmaf(cosVal, bp_cell_sum(pp, u0, nu, wu0, fh, vlo, vhi, DNV), acc[g]);
      }
    }
  }
#pragma unroll
  for (int g = 0; g < ZG; ++g) {
    const int z = zb + 64 * g;
    if (z < d.ZN) bp_store<UPDATE>(A, vcol, col, z, acc[g]);
  }
}

// Windowed form of bp_ref_body: one wave per voxel column over all of its z.  On a
// helical scan a view's cone covers a window of a few dozen voxels of a column, which
// the fixed lane -> z groups above cover with one 64-lane pass per group it touches
// (about half the lanes idle).  Here the lanes take the view's own window, z0 + lane,
// z0 + 64 + lane, ... (z0, z1: the voxels of the row shadow's open interval, with a
// 1e-2-voxel margin; the per-voxel test decides), and the column's sums live in LDS
// (lacc: ZN floats of this wave), each
// updated by one lane per view, views in order.  Per-voxel arithmetic as bp_ref_body.
// ONES: the back projection of a sinogram of ones (the column sums of OS-SART), read
// from no sinogram: the row weights of a shadow sum to its clipped height vhi - vlo and
// the column weights to Wu = wu0 + (nu - 2) + fh (wu0 alone for one column), so a voxel
// adds cosVal * Wu * (vhi - vlo) (the same sum in another rounding order).
//
// stg (kStageF floats of this wave, 16-byte aligned, or NULL): with DNV <= kStageRows the
// views are staged (lacc then holds ZN + kLaccPad floats).  Lane l writes view k0 + l's record (row-shadow and cosVal constants,
// column weights) to the wave's table; then, visible view by visible view, every lane
// loads its row of the view's first four columns (issued while the previous view's voxels
// are summed), combines them (R[v] = sum_u wu y[u][v], the column order and rounding of
// bp_cells) into LDS, and the voxels of the window read two or three values of R there;
// the sums are those of the direct gathers, bit for bit.
template <bool UPDATE, typename T, bool ONES = false>
__device__ __forceinline__ void bp_win_body(const BpArgs<T>& A, int col, float* lacc,
                                            float* stg) {
  const Dims& d = A.d;
  const BpConsts64& c64 = A.c64;
  const int lane = threadIdx.x & 63;
  if (col >= d.XN * d.YN) return;
  const int y = col / d.XN, x = col - y * d.XN;
  CTR_GLOBAL float* vcol = gp(A.vol) + (size_t)col * d.ZN;
  if (gp(A.mask)[col] != 1) {
    if constexpr (!UPDATE) {
      for (int z = lane; z < d.ZN; z += 64) vcol[z] = 0.f;
    }
    return;
  }
  for (int z = lane; z < d.ZN; z += 64) lacc[z] = 0.f;
  const double Px = ((double)x - c64.ocx) * c64.dx;
  const double Py = ((double)y - c64.ocy) * c64.dx;
  const float Pxf = (float)Px, Pyf = (float)Py;
  const float DNVh = (float)d.DNV - 0.5f;
  const float ZNf = (float)d.ZN;
  const float idz = 1.f / d.dz;
  const int DNV = d.DNV;
  const uint32_t frame = (uint32_t)(d.DNU * d.DNV);
  const bool stage = !ONES && stg != nullptr && DNV <= kStageRows;
  float* lR = stg;                                         // rows of the current view
  float4* rec = reinterpret_cast<float4*>(stg + kRowsF);  // [64][3] float4 records
  if (stage && lane < kStagePad - kStageRows) lR[kStageRows + lane] = 0.f;
  const int row = min(lane, DNV - 1);  // the row a lane stages
  const float zlane = (float)lane - d.ocz;
  for (int k0 = 0; k0 < A.vc; k0 += 64) {
    const int kl = k0 + lane;
    BpView q{};
    int q_z0 = 0, q_z1 = -1;
    if (kl < A.vc) {
      q = bp_view(A, A.vb + kl * A.vs, Px, Py, Pxf, Pyf);
      // voxels with fmaf(dZ, rlS, cB) > -0.5 and fmaf(dZ, rlS, cA) < DNV - 0.5 lie in the
      // open interval (za, zb) of voxel indices; za, zb are within ~1e-4 of it (FP32), so
      // the integers inside (za - m, zb + m), m = 1e-2, include every one of them
      const float ea = (-0.5f - q.cB) / q.rlS, eb = (DNVh - q.cA) / q.rlS;
      const float za = fminf(fmaxf((fminf(ea, eb) + q.sz) * idz + d.ocz, -4.f), ZNf + 4.f);
      const float zb = fminf(fmaxf((fmaxf(ea, eb) + q.sz) * idz + d.ocz, -4.f), ZNf + 4.f);
      q_z0 = max((int)floorf(za - 1e-2f) + 1, 0);
      q_z1 = min((int)ceilf(zb + 1e-2f) - 1, d.ZN - 1);
    }
    uint64_t live = __ballot(q.uvis && q_z0 <= q_z1);
    if (stage && live) {
#pragma clang fp contract(off)
      {
        // column weights of view k0 + lane: first wu0, interior 1, last fh, past nu 0
        const int nu = q.nu;
        const float w1 = nu > 1 ? (nu == 2 ? q.fh : 1.f) : 0.f;
        const float w2 = nu > 2 ? (nu == 3 ? q.fh : 1.f) : 0.f;
        const float w3 = nu > 3 ? (nu == 4 ? q.fh : 1.f) : 0.f;
        rec[3 * lane + 0] = make_float4(q.rlS, q.cA, q.cB, q.sz);
        rec[3 * lane + 1] = make_float4(q.l2, q.cws, q.wu0, w1);
        rec[3 * lane + 2] = make_float4(w2, w3, 0.f, 0.f);
      }
      // rows of view k0 + j: columns u0 + min(iu, nu - 1) (a repeated column has weight 0)
      // (stored bits, widened at their use: the next view's loads stay in flight while
      // this view's voxels are summed)
      auto load_rows = [&](int j, uint32_t (&yb)[4]) {
        const int un = __builtin_amdgcn_readlane(q.u0, j);
        const int nn = __builtin_amdgcn_readlane(q.nu, j);
        const Buf<T> pn(A.prj + (size_t)(k0 + j) * frame, frame);
#pragma unroll
        for (int iu = 0; iu < 4; ++iu) yb[iu] = pn.at_bits(row, (un + min(iu, nn - 1)) * DNV);
      };
      int j = __builtin_ctzll(live);
      uint32_t yb[4];
      load_rows(j, yb);
      while (true) {
        const int nu = __builtin_amdgcn_readlane(q.nu, j);
        const int z0 = __builtin_amdgcn_readlane(q_z0, j), z1 = __builtin_amdgcn_readlane(q_z1, j);
        const float4 r0 = rec[3 * j], r1 = rec[3 * j + 1], r2 = rec[3 * j + 2];
        float R = r1.z * Buf<T>::widen(yb[0]);
        R = fmaf(r1.w, Buf<T>::widen(yb[1]), R);
        R = fmaf(r2.x, Buf<T>::widen(yb[2]), R);
        R = fmaf(r2.y, Buf<T>::widen(yb[3]), R);
        if (nu > 4) {
          const int un = __builtin_amdgcn_readlane(q.u0, j);
          const float fh = rdl(q.fh, j);
          const Buf<T> pn(A.prj + (size_t)(k0 + j) * frame, frame);
          for (int iu = 4; iu < nu; ++iu)
            R = fmaf((iu == nu - 1) ? fh : 1.f, pn.at(row, (un + iu) * DNV), R);
        }
        lR[lane] = lane < DNV ? R : 0.f;
        // the next visible view's rows load while this view's voxels are summed
        live &= live - 1;
        const int jn = live ? __builtin_ctzll(live) : -1;
        if (jn >= 0) load_rows(jn, yb);
        const float rlS = r0.x, cA = r0.y, cB = r0.z, vsz = r0.w, l2f = r1.x, cws = r1.y;
        for (int zz = z0; zz <= z1; zz += 64) {
          // (float)(zz + lane) - ocz, exactly (half-integers far below 2^23); voxels past
          // z1 lie outside the shadow's open interval with the window's 1e-2-voxel margin
          // (empty shadows), those past the column add to lacc's padding
          const float dZ = fmaf((float)zz + zlane, d.dz, -vsz);
          const float vlo = fmaxf(fmaf(dZ, rlS, cA), -0.5f);
          const float vhi = __builtin_elementwise_minimum(fmaf(dZ, rlS, cB), DNVh);
          if (vlo < vhi) {
            const float cosVal = cws * __builtin_amdgcn_sqrtf(fmaf(dZ, dZ, l2f));
            float* lz = lacc + zz;
            lz[lane] = fmaf(cosVal, bp_lds_sum(lR, vlo, vhi, DNV), lz[lane]);
          }
        }
        if (jn < 0) break;
        j = jn;
      }
      continue;
    }
    if constexpr (ONES) {
      // views' constants (with the column-weight sum Wu) from a table in LDS (stg: 64
      // records of 8 floats), read at a uniform address instead of by readlanes
      if (live) {
        float4* rec1 = reinterpret_cast<float4*>(stg);
        const int nu = q.nu;
        const float Wu = nu == 1 ? q.wu0 : q.wu0 + (float)(nu - 2) + q.fh;
        rec1[2 * lane] = make_float4(q.rlS, q.cA, q.cB, q.sz);
        rec1[2 * lane + 1] = make_float4(q.l2, q.cws, Wu, 0.f);
        while (live) {
          const int j = __builtin_ctzll(live);
          live &= live - 1;
          const int z0 = __builtin_amdgcn_readlane(q_z0, j), z1 = __builtin_amdgcn_readlane(q_z1, j);
          const float4 r0 = rec1[2 * j], r1 = rec1[2 * j + 1];
          const float rlS = r0.x, cA = r0.y, cB = r0.z, vsz = r0.w;
          const float l2f = r1.x, cws = r1.y, Wv = r1.z;
          for (int zz = z0; zz <= z1; zz += 64) {
            const int z = zz + lane;
            const float dZ = fmaf((float)z - d.ocz, d.dz, -vsz);
            const float vlo = fmaxf(fmaf(dZ, rlS, cA), -0.5f);
            const float vhi = fminf(fmaf(dZ, rlS, cB), DNVh);
            if (z <= z1 && vlo < vhi) {
              const float cosVal = cws * __builtin_amdgcn_sqrtf(fmaf(dZ, dZ, l2f));
              lacc[z] = fmaf(cosVal, Wv * (vhi - vlo), lacc[z]);
            }
          }
        }
      }
      continue;
    }
    while (live) {
      const int j = __builtin_ctzll(live);
      live &= live - 1;
      const int k = k0 + j;
      const float rlS = rdl(q.rlS, j), cA = rdl(q.cA, j), cB = rdl(q.cB, j);
      const float vsz = rdl(q.sz, j), l2f = rdl(q.l2, j), cws = rdl(q.cws, j);
      const float wu0 = rdl(q.wu0, j), fh = rdl(q.fh, j);
      const int u0 = __builtin_amdgcn_readlane(q.u0, j), nu = __builtin_amdgcn_readlane(q.nu, j);
      const int z0 = __builtin_amdgcn_readlane(q_z0, j), z1 = __builtin_amdgcn_readlane(q_z1, j);
      const Buf<T> pp(A.prj + (size_t)k * frame, frame);
      const float Wu = nu == 1 ? wu0 : wu0 + (float)(nu - 2) + fh;
      for (int zz = z0; zz <= z1; zz += 64) {
        const int z = zz + lane;
        const float dZ = fmaf((float)z - d.ocz, d.dz, -vsz);
        const float vlo = fmaxf(fmaf(dZ, rlS, cA), -0.5f);
        const float vhi = fminf(fmaf(dZ, rlS, cB), DNVh);
        const bool vok = z <= z1 && vlo < vhi;
        if (!vok) continue;
        const float cosVal = cws * __builtin_amdgcn_sqrtf(fmaf(dZ, dZ, l2f));
        if constexpr (ONES)
          lacc[z] = fmaf(cosVal, Wu * (vhi - vlo), lacc[z]);
        else
          lacc[z] = fmaf(cosVal, bp_cell_sum(pp, u0, nu, wu0, fh, vlo, vhi, DNV), lacc[z]);
      }
    }
  }
  for (int z = lane; z < d.ZN; z += 64) bp_store<UPDATE>(A, vcol, col, z, lacc[z]);
}

template <bool UPDATE, int ZG, typename T>
__global__ __launch_bounds__(kBlock) void bp_ref_kernel(BpArgs<T> A) {
  const int col = blockIdx.x * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  bp_ref_body<UPDATE, ZG, T>(A, col, blockIdx.y);
}

// Windowed reference back projection; dynamic LDS: kWaves * kStageF floats of staged
// views (DNV <= kStageRows; none otherwise), then kWaves * (ZN + kLaccPad) floats of
// column sums (kWaves * ZN unstaged).
// At most 72 VGPRs (7 waves per SIMD; no scratch).
template <bool UPDATE, typename T>
__global__ __launch_bounds__(kBlock, 7) void bp_win_kernel(BpArgs<T> A) {
  extern __shared__ float4 lds_win[];
  float* lds = reinterpret_cast<float*>(lds_win);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool stage = A.d.DNV <= kStageRows;
  float* lacc = lds + (stage ? kWaves * kStageF : 0) + wid * (A.d.ZN + (stage ? kLaccPad : 0));
  bp_win_body<UPDATE, T>(A, blockIdx.x * kWaves + wid, lacc,
                         stage ? lds + wid * kStageF : nullptr);
}

// Back projection of a sinogram of ones (A.prj unused); dynamic LDS: kWaves * 512
// floats of view records, then kWaves * ZN floats of column sums.
constexpr int kOnesRecF = 512;
__global__ __launch_bounds__(kBlock, 8) void bp_win_ones_kernel(BpArgs<float> A) {
  extern __shared__ float4 lds_ones[];
  float* lds = reinterpret_cast<float*>(lds_ones);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  bp_win_body<false, float, true>(A, blockIdx.x * kWaves + wid,
                                  lds + kWaves * kOnesRecF + wid * A.d.ZN,
                                  lds + wid * kOnesRecF);
}

// Batched windowed reference back projection: flat work list of (scan, block of 4
// columns); dynamic LDS: kWaves * kStageF floats of staged views (every scan of the batch
// has DNV <= kStageRows: staged; otherwise none), then kWaves * (zmax + kLaccPad) floats
// (zmax: the tallest column of the batch; kWaves * zmax unstaged).
template <typename T>
__global__ __launch_bounds__(kBlock, 7) void bp_batch_win_kernel(const BpBatchItem<T>* __restrict__ D,
                                                              int nscan, int total, int sched,
                                                              int zmax, int staged) {
  extern __shared__ float4 lds_win[];
  float* lds = reinterpret_cast<float*>(lds_win);
  int i0, step, i1;
  batch_range(D, nscan, total, sched, i0, step, i1);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int s = 0;
  for (int item = i0; item < i1; item += step) {
    if (sched == 1) s = 0;
    while (s + 1 < nscan && D[s + 1].begin <= item) ++s;
    const BpBatchItem<T>& B = D[s];
    bp_win_body<true, T>(B.a, (item - B.begin) * kWaves + wid,
                         lds + (staged ? kWaves * kStageF : 0) +
                             wid * (zmax + (staged ? kLaccPad : 0)),
                         staged ? lds + wid * kStageF : nullptr);
  }
}

// bp_batch_win_kernel for the one-workgroup-per-item schedule (grid = total): no item
// loop, so nothing stays live around the column body.  Both batched windowed kernels are
// held to 72 VGPRs (7 waves per SIMD), as bp_win_kernel.
template <typename T>
__global__ __launch_bounds__(kBlock, 7) void bp_batch_win1_kernel(
    const BpBatchItem<T>* __restrict__ D, int nscan, int total, int zmax, int staged) {
  extern __shared__ float4 lds_win[];
  float* lds = reinterpret_cast<float*>(lds_win);
  int i0, step, i1;
  batch_range(D, nscan, total, 2, i0, step, i1);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int s = 0;
  while (s + 1 < nscan && D[s + 1].begin <= i0) ++s;
  const BpBatchItem<T>& B = D[s];
  bp_win_body<true, T>(B.a, (i0 - B.begin) * kWaves + wid,
                       lds + (staged ? kWaves * kStageF : 0) +
                           wid * (zmax + (staged ? kLaccPad : 0)),
                       staged ? lds + wid * kStageF : nullptr);
}

// Batched persistent reference back projection: flat work list of (scan, z block,
// block of 4 columns), contiguous ranges per workgroup as in fp_batch_kernel.
template <bool UPDATE, int ZG, typename T>
__global__ __launch_bounds__(kBlock) void bp_batch_kernel(const BpBatchItem<T>* __restrict__ D,
                                                          int nscan, int total, int sched) {
  int i0, step, i1;
  batch_range(D, nscan, total, sched, i0, step, i1);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int s = 0;
  for (int item = i0; item < i1; item += step) {
    if (sched == 1) s = 0;
    while (s + 1 < nscan && D[s + 1].begin <= item) ++s;
    const BpBatchItem<T>& B = D[s];
    const int local = item - B.begin;
    const int zblk = local / B.colBlocks;
    bp_ref_body<UPDATE, ZG, T>(B.a, (local - zblk * B.colBlocks) * kWaves + wid, zblk);
  }
}


// ---------------------------------------------------------------------------
// Matched adjoint: the exact transpose of fp_kernel, in gather form.
// Input yp is the sinogram already multiplied by K * invdB (scale_matched_kernel).
// For each view the voxel's in-plane cell is located among the column footprints
// on its slice with the FP64 boundary lines fp_kernel evaluates, and its z cell
// among the row footprints (FP32, relative to the view's azi as in fp_kernel); the
// overlap weights are those of the forward projector.
// ---------------------------------------------------------------------------
template <bool UPDATE, int ZG>
__global__ __launch_bounds__(kBlock) void bp_matched_kernel(
    const float* __restrict__ yp, float* __restrict__ vol, const ViewP* __restrict__ views,
    const double2* __restrict__ bnd, const float4* __restrict__ cols,
    const float* __restrict__ Bz, const uint8_t* __restrict__ mask, Dims d, BpConsts c,
    int vb, int vs, int vc, const float* __restrict__ invCol, uint16_t* __restrict__ shadow) {
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (col >= d.XN * d.YN) return;
  const int zb = blockIdx.y * (64 * ZG) + lane;
  const int y = col / d.XN, x = col - y * d.XN;
  float* vcol = vol + (size_t)col * d.ZN;
  const float m = (float)mask[col];
  if (m == 0.f) {
    if constexpr (!UPDATE) {
#pragma unroll
      for (int g = 0; g < ZG; ++g)
        if (zb + 64 * g < d.ZN) vcol[zb + 64 * g] = 0.f;
    }
    return;
  }
  const float Px = ((float)x - d.ocx) * d.dx;
  const float Py = ((float)y - d.ocy) * d.dx;
  float acc[ZG];
  bool live[ZG];
#pragma unroll
  for (int g = 0; g < ZG; ++g) {
    acc[g] = 0.f;
    live[g] = zb + 64 * g < d.ZN;
  }
  const size_t frame = (size_t)d.DNU * d.DNV;
  const int DNU = d.DNU, DNV = d.DNV;
  for (int k = 0; k < vc; ++k) {
    const int p = vb + k * vs;
    const ViewP vp = views[p];
    const int ii = vp.xmarch ? x : y;
    const int j = vp.xmarch ? y : x;
    const double fi = (double)ii;
    const double2* bt = bnd + (size_t)p * (DNU + 1);
    const float4* ctab = cols + (size_t)p * DNU;
    const double dir = (double)vp.bdir;
    const double cl = (vp.bdir > 0) ? (double)j - 0.5 : -((double)j + 0.5);
    const double ch = cl + 1.0;
    // first guess from the fan angle of the voxel centre, then walk to the first
    // column whose footprint (the FP64 boundary lines of fp_kernel) reaches cl
    const float dX = Px - vp.sx, dY = Py - vp.sy;
    const float beta = asinf((vp.sy * dX - vp.sx * dY) * c.inv_sid * rsqrtf(dX * dX + dY * dY));
    int u = (int)fminf(fmaxf(floorf((beta - d.beta0) / d.dbeta_g), 0.f), (float)(DNU - 1));
    double su = dir * fma(fi, bt[u].y, bt[u].x);
    while (u > 0 && su > cl) {
      --u;
      su = dir * fma(fi, bt[u].y, bt[u].x);
    }
    double su1 = dir * fma(fi, bt[u + 1].y, bt[u + 1].x);
    while (u < DNU - 1 && su1 <= cl) {
      ++u;
      su = su1;
      su1 = dir * fma(fi, bt[u + 1].y, bt[u + 1].x);
    }
    const float* yk = yp + (size_t)k * frame;
    // z relative to azi, as fp_kernel
    const float az = vp.azf;
    for (; u < DNU && su < ch;) {
      const double wy = fmin(su1, ch) - fmax(su, cl);
      if (wy > 0.0) {
        const float4 ct = ctab[u];
        const float r = fmaf((float)ii, ct.y, ct.x);
        const float wyi = (float)wy * __builtin_amdgcn_rcpf((float)(su1 - su) * r);
        const float* yu = yk + (size_t)u * DNV;
        const float rinv = 1.0f / r;
        const float zlo0 = fmaf(r, Bz[0], az), zhiN = fmaf(r, Bz[DNV], az);
#pragma unroll
        for (int g = 0; g < ZG; ++g) {
          // rows whose z footprint [az + Bz[v] r, az + Bz[v+1] r] meets [zlo, zhi]
          const float zlo = (float)(zb + 64 * g - vp.azi) - 0.5f, zhi = zlo + 1.0f;
          const bool vis = live[g] && zhi > zlo0 && zlo < zhiN;
          if (!__any(vis)) continue;
          if (!vis) continue;
          int v = (int)fminf(fmaxf(floorf(((zlo - az) * rinv - d.B0) / d.dB), 0.f),
                             (float)(DNV - 1));
          float zl = fmaf(r, Bz[v], az);
          while (v > 0 && zl > zlo) {
            --v;
            zl = fmaf(r, Bz[v], az);
          }
          float zh = fmaf(r, Bz[v + 1], az);
          while (v < DNV - 1 && zh <= zlo) {
            ++v;
            zl = zh;
            zh = fmaf(r, Bz[v + 1], az);
          }
          float gsum = 0.f;
          for (; v < DNV && zl < zhi;) {
            const float wz = fminf(zh, zhi) - fmaxf(zl, zlo);
            if (wz > 0.f) gsum = fmaf(wz, yu[v], gsum);
            ++v;
            if (v < DNV) {
              zl = zh;
              zh = fmaf(r, Bz[v + 1], az);
            }
          }
          acc[g] = fmaf(wyi, gsum, acc[g]);
        }
      }
      ++u;
      if (u < DNU) {
        su = su1;
        su1 = dir * fma(fi, bt[u + 1].y, bt[u + 1].x);
      }
    }
  }
#pragma unroll
  for (int g = 0; g < ZG; ++g) {
    if (!live[g]) continue;
    const int z = zb + 64 * g;
    const float a = acc[g] * m;
    if constexpr (UPDATE) {
      const float nx = fmaf(a, invCol[(size_t)col * d.ZN + z], vcol[z]);
      vcol[z] = nx;
      if (shadow) shadow[(size_t)col * d.ZN + z] = to_bf16_dither(nx, (size_t)col * d.ZN + z);
    } else {
      vcol[z] = a;
    }
  }
}

// Rows of one detector column meeting the z cell [zlo, zlo + 1] of a voxel, integrated
// against the column's scaled sinogram yu with the forward projector's FP32 row
// boundaries az + r Bz[v] (Bz from LDS); the first row from the uniform-spacing guess,
// then corrected against the exact boundaries.
__device__ __forceinline__ float matched_rows(const float* __restrict__ yu,
                                              const float* __restrict__ sBz, int DNV, float r,
                                              float az, float zlo, float B0, float dB) {
  const float zhi = zlo + 1.0f;
  int v = (int)fminf(fmaxf(floorf(((zlo - az) * __builtin_amdgcn_rcpf(r) - B0) / dB), 0.f),
                     (float)(DNV - 1));
  float zl = fmaf(r, sBz[v], az);
  while (v > 0 && zl > zlo) {
    --v;
    zl = fmaf(r, sBz[v], az);
  }
  float zh = fmaf(r, sBz[v + 1], az);
  while (v < DNV - 1 && zh <= zlo) {
    ++v;
    zl = zh;
    zh = fmaf(r, sBz[v + 1], az);
  }
  float gsum = 0.f;
  for (; v < DNV && zl < zhi;) {
    const float wz = fminf(zh, zhi) - fmaxf(zl, zlo);
    if (wz > 0.f) gsum = fmaf(wz, yu[v], gsum);
    ++v;
    if (v < DNV) {
      zl = zh;
      zh = fmaf(r, sBz[v + 1], az);
    }
  }
  return gsum;
}

// The same sums as matched_rows for the (at most four) columns t < nu of one view, the
// rows taken as a fixed count NR (at least the number of rows of spacing r_t dBmin that
// can meet a cell of height 1): every column's first row is searched, then every
// sinogram value is loaded, then the rows are summed in order, rows with no overlap or
// past the detector adding nothing -- the sums and their order are those of matched_rows.
// ig[t] = 1 / (r_t dB), g0 = -B0 / dB give the first-row guess.
template <int NR>
__device__ __forceinline__ float matched_cols(const float* __restrict__ yk,
                                              const float* __restrict__ sBz, int DNV, int nu,
                                              const float (&w)[4], const float (&r)[4],
                                              const float (&ig)[4], float az, float g0,
                                              float zlo, float a) {
  const float zhi = zlo + 1.0f;
  int v[4];
  float zl[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    v[t] = 0;
    zl[t] = 0.f;
    if (t < nu) {
      int vv = (int)fminf(fmaxf(floorf(fmaf(zlo - az, ig[t], g0)), 0.f), (float)(DNV - 1));
      float l = fmaf(r[t], sBz[vv], az);
      while (vv > 0 && l > zlo) {
        --vv;
        l = fmaf(r[t], sBz[vv], az);
      }
      float h = fmaf(r[t], sBz[vv + 1], az);
      while (vv < DNV - 1 && h <= zlo) {
        ++vv;
        l = h;
        h = fmaf(r[t], sBz[vv + 1], az);
      }
      v[t] = vv;
      zl[t] = l;
    }
  }
  float yv[4][NR];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int i = 0; i < NR; ++i)
      yv[t][i] = (t < nu) ? yk[(size_t)t * DNV + min(v[t] + i, DNV - 1)] : 0.f;
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if (t < nu && w[t] != 0.f) {
      float g = 0.f, lo = zl[t];
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        const int vi = v[t] + i;
        const float hi = fmaf(r[t], sBz[min(vi + 1, DNV)], az);
        const float wz = fminf(hi, zhi) - fmaxf(lo, zlo);
        if (vi < DNV && wz > 0.f) g = fmaf(wz, yv[t][i], g);
        lo = hi;
      }
      a = fmaf(w[t], g, a);
    }
  }
  return a;
}

// Windowed matched adjoint: the transpose of fp_kernel as bp_matched_kernel computes it,
// with one wave per voxel column over all of its z.  Per view (64 views at a time, one
// per lane): the detector columns whose footprints on the voxel's slice meet its cell
// (first column u0, count nu; for nu <= 4 the in-plane weights w_t = overlap / ((tr - tl)
// r) and magnifications r_t), and the view's z-window of the column (the cells the
// detector's z range at those magnifications reaches, widened by one).  The lanes then
// take the window, each voxel summing w_t times its rows' integral over the columns in
// order (the same sequence of operations per voxel as bp_matched_kernel), into the
// column's LDS sums.  Dynamic LDS: Bz (DNV + 1 floats, 4-aligned), then kWaves * ZN.
template <bool UPDATE>
__global__ __launch_bounds__(kBlock) void bp_matched_win_kernel(
    const float* __restrict__ yp, float* __restrict__ vol, const ViewP* __restrict__ views,
    const double2* __restrict__ bnd, const float4* __restrict__ cols,
    const float* __restrict__ Bz, const uint8_t* __restrict__ mask, Dims d, BpConsts c,
    int vb, int vs, int vc, const float* __restrict__ invCol, uint16_t* __restrict__ shadow) {
  extern __shared__ float lds_m[];
  const int DNU = d.DNU, DNV = d.DNV;
  const int nbz = (DNV + 1 + 3) & ~3;
  float* sBz = lds_m;
  for (int i = threadIdx.x; i <= DNV; i += kBlock) sBz[i] = Bz[i];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col = blockIdx.x * kWaves + wid;
  if (col >= d.XN * d.YN) return;
  float* lacc = lds_m + nbz + wid * d.ZN;
  const int y = col / d.XN, x = col - y * d.XN;
  float* vcol = vol + (size_t)col * d.ZN;
  const float m = (float)mask[col];
  if (m == 0.f) {
    if constexpr (!UPDATE) {
      for (int z = lane; z < d.ZN; z += 64) vcol[z] = 0.f;
    }
    return;
  }
  for (int z = lane; z < d.ZN; z += 64) lacc[z] = 0.f;
  const float Px = ((float)x - d.ocx) * d.dx;
  const float Py = ((float)y - d.ocy) * d.dx;
  const size_t frame = (size_t)DNU * DNV;
  const float ZNf = (float)d.ZN;
  for (int k0 = 0; k0 < vc; k0 += 64) {
    const int kl = k0 + lane;
    int q_u0 = 0, q_nu = 0, q_z0 = 0, q_z1 = -1;
    float q_w[4] = {0.f, 0.f, 0.f, 0.f}, q_r[4] = {1.f, 1.f, 1.f, 1.f};
    float q_ig[4] = {0.f, 0.f, 0.f, 0.f};
    float q_az = 0.f;
    int q_azi = 0;
    int q_nr = 0;
    if (kl < vc) {
      const int p = vb + kl * vs;
      const ViewP vp = views[p];
      const int ii = vp.xmarch ? x : y;
      const int j = vp.xmarch ? y : x;
      const double fi = (double)ii;
      const double2* bt = bnd + (size_t)p * (DNU + 1);
      const float4* ctab = cols + (size_t)p * DNU;
      const double dir = (double)vp.bdir;
      const double cl = (vp.bdir > 0) ? (double)j - 0.5 : -((double)j + 0.5);
      const double ch = cl + 1.0;
      // first guess from the fan angle of the voxel centre, then walk to the first
      // column whose footprint (the FP64 boundary lines of fp_kernel) reaches cl
      const float dX = Px - vp.sx, dY = Py - vp.sy;
      const float beta =
          asinf((vp.sy * dX - vp.sx * dY) * c.inv_sid * rsqrtf(dX * dX + dY * dY));
      int u = (int)fminf(fmaxf(floorf((beta - d.beta0) / d.dbeta_g), 0.f), (float)(DNU - 1));
      double su = dir * fma(fi, bt[u].y, bt[u].x);
      while (u > 0 && su > cl) {
        --u;
        su = dir * fma(fi, bt[u].y, bt[u].x);
      }
      double su1 = dir * fma(fi, bt[u + 1].y, bt[u + 1].x);
      while (u < DNU - 1 && su1 <= cl) {
        ++u;
        su = su1;
        su1 = dir * fma(fi, bt[u + 1].y, bt[u + 1].x);
      }
      // the columns u .. while su < ch (zero-overlap ones included: they add nothing)
      q_u0 = u;
      float rmin = 3.4e38f, rmax = -3.4e38f;
      int n = 0;
      for (; u < DNU && su < ch;) {
        const float4 ct = ctab[u];
        const float r = fmaf((float)ii, ct.y, ct.x);
        const double wy = fmin(su1, ch) - fmax(su, cl);
        if (n < 4) {
          q_w[n] = wy > 0.0 ? (float)wy * __builtin_amdgcn_rcpf((float)(su1 - su) * r) : 0.f;
          q_r[n] = r;
          q_ig[n] = __builtin_amdgcn_rcpf(r * d.dB);
        }
        rmin = fminf(rmin, r);
        rmax = fmaxf(rmax, r);
        ++n;
        ++u;
        if (u < DNU) {
          su = su1;
          su1 = dir * fma(fi, bt[u + 1].y, bt[u + 1].x);
        }
      }
      q_nu = n;
      // z relative to azi, as fp_kernel
      q_az = vp.azf;
      q_azi = vp.azi;
      if (n > 0) {
        // z cells [k - .5, k + .5] meeting [az + r Bz[0], az + r Bz[DNV]] for some r of
        // the columns (r > 0 beyond the source)
        const float lo = fminf(fmaf(rmin, Bz[0], vp.azf), fmaf(rmax, Bz[0], vp.azf));
        const float hi = fmaxf(fmaf(rmin, Bz[DNV], vp.azf), fmaf(rmax, Bz[DNV], vp.azf));
        const float fazi = (float)vp.azi;
        const float za = fminf(fmaxf(lo - 0.5f, -4.f - fazi), (ZNf + 4.f) - fazi);
        const float zb = fminf(fmaxf(hi + 0.5f, -4.f - fazi), (ZNf + 4.f) - fazi);
        q_z0 = max((int)floorf(za) - 1 + vp.azi, 0);
        q_z1 = min((int)ceilf(zb) + 1 + vp.azi, d.ZN - 1);
        // rows of spacing >= h meeting a cell of height 1: at most ceil(1 / h) + 1; h is
        // the smallest spacing rmin dBmin less the rounding of boundaries of size <= big
        const float big = fmaxf(fmaxf(fabsf(lo), fabsf(hi)), 1.f);
        const float h = rmin * d.dBmin * 0.99999f - 4.8e-7f * big;
        q_nr = h > 0.125f ? (int)ceilf(__builtin_amdgcn_rcpf(h) * 1.0001f) + 1 : 99;
      }
    }
    uint64_t live = __ballot(q_nu > 0 && q_z0 <= q_z1);
    while (live) {
      const int jl = __builtin_ctzll(live);
      live &= live - 1;
      const int k = k0 + jl;
      const int u0 = __builtin_amdgcn_readlane(q_u0, jl), nu = __builtin_amdgcn_readlane(q_nu, jl);
      const int z0 = __builtin_amdgcn_readlane(q_z0, jl), z1 = __builtin_amdgcn_readlane(q_z1, jl);
      const float az = rdl(q_az, jl);
      const int azi = __builtin_amdgcn_readlane(q_azi, jl);
      const int nr = __builtin_amdgcn_readlane(q_nr, jl);
      const float* yk = yp + (size_t)k * frame + (size_t)u0 * DNV;
      if (nu <= 4 && nr <= 5) {
        float w[4], r[4], ig[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          w[t] = rdl(q_w[t], jl);
          r[t] = rdl(q_r[t], jl);
          ig[t] = rdl(q_ig[t], jl);
        }
        const float g0 = -d.B0 / d.dB;
        for (int zz = z0; zz <= z1; zz += 64) {
          const int z = zz + lane;
          if (z > z1) continue;
          const float zlo = (float)(z - azi) - 0.5f;
          const float a0 = lacc[z];
          float a;
          switch (nr) {
            case 2: a = matched_cols<2>(yk, sBz, DNV, nu, w, r, ig, az, g0, zlo, a0); break;
            case 3: a = matched_cols<3>(yk, sBz, DNV, nu, w, r, ig, az, g0, zlo, a0); break;
            case 4: a = matched_cols<4>(yk, sBz, DNV, nu, w, r, ig, az, g0, zlo, a0); break;
            default: a = matched_cols<5>(yk, sBz, DNV, nu, w, r, ig, az, g0, zlo, a0); break;
          }
          lacc[z] = a;
        }
      } else if (nu <= 4) {
        float w[4], r[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          w[t] = rdl(q_w[t], jl);
          r[t] = rdl(q_r[t], jl);
        }
        for (int zz = z0; zz <= z1; zz += 64) {
          const int z = zz + lane;
          if (z > z1) continue;
          const float zlo = (float)(z - azi) - 0.5f;
          float a = lacc[z];
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            if (t < nu && w[t] != 0.f)
              a = fmaf(w[t], matched_rows(yk + t * DNV, sBz, DNV, r[t], az, zlo, d.B0, d.dB),
                       a);
          }
          lacc[z] = a;
        }
      } else {
        // more than four columns: their weights recomputed here, as in the lane stage
        const int p = vb + k * vs;
        const ViewP vp = views[p];
        const int ii = vp.xmarch ? x : y;
        const int j = vp.xmarch ? y : x;
        const double fi = (double)ii;
        const double2* bt = bnd + (size_t)p * (DNU + 1);
        const float4* ctab = cols + (size_t)p * DNU;
        const double dir = (double)vp.bdir;
        const double cl = (vp.bdir > 0) ? (double)j - 0.5 : -((double)j + 0.5);
        const double ch = cl + 1.0;
        for (int zz = z0; zz <= z1; zz += 64) {
          const int z = zz + lane;
          if (z > z1) continue;
          const float zlo = (float)(z - azi) - 0.5f;
          float a = lacc[z];
          for (int t = 0; t < nu; ++t) {
            const int u = u0 + t;
            const double su = dir * fma(fi, bt[u].y, bt[u].x);
            const double su1 = dir * fma(fi, bt[u + 1].y, bt[u + 1].x);
            const double wy = fmin(su1, ch) - fmax(su, cl);
            if (wy > 0.0) {
              const float4 ct = ctab[u];
              const float r = fmaf((float)ii, ct.y, ct.x);
              const float wyi = (float)wy * __builtin_amdgcn_rcpf((float)(su1 - su) * r);
              a = fmaf(wyi, matched_rows(yk + t * DNV, sBz, DNV, r, az, zlo, d.B0, d.dB), a);
            }
          }
          lacc[z] = a;
        }
      }
    }
  }
  for (int z = lane; z < d.ZN; z += 64) {
    const float a = lacc[z] * m;
    if constexpr (UPDATE) {
      const size_t i = (size_t)col * d.ZN + z;
      const float nx = fmaf(a, invCol[i], vcol[z]);
      vcol[z] = nx;
      if (shadow) shadow[i] = to_bf16_dither(nx, i);
    } else {
      vcol[z] = a;
    }
  }
}

// yp[k][u][v] = y[k][u][v] * K(p,u,v) * invdB[v]: the per-ray factors of fp_kernel
// (y stored as float or BF16, yp FP32).
template <typename T>
__global__ __launch_bounds__(kBlock) void scale_matched_kernel(
    const T* __restrict__ in, float* __restrict__ out, const float4* __restrict__ cols,
    const float* __restrict__ h2, const float* __restrict__ invdB, Dims d, int vb, int vs) {
  const int k = blockIdx.y;
  const int p = vb + k * vs;
  const int ray = blockIdx.x * kBlock + threadIdx.x;
  if (ray >= d.DNU * d.DNV) return;
  const int u = ray / d.DNV, v = ray - u * d.DNV;
  const float4 ct = cols[(size_t)p * d.DNU + u];
  const float K = ct.w * sqrtf(ct.z + h2[v]) * invdB[v];
  const size_t o = (size_t)k * d.DNU * d.DNV + ray;
  out[o] = ld(in, o) * K;
}

__global__ void copy_kernel(float* __restrict__ dst, const float* __restrict__ src, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void fill_kernel(float* __restrict__ p, size_t n, float v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void to_bf16_kernel(uint16_t* __restrict__ dst, const float* __restrict__ src,
                               size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = to_bf16(src[i]);
}

__global__ void to_bf16_dither_kernel(uint16_t* __restrict__ dst, const float* __restrict__ src,
                                      size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = to_bf16_dither(src[i], i);
}

// vol[col][z] *= mask[col] (the in-place masking of DD3_GPU.cu:1004-1011, on the device).
__global__ void mask_volume_kernel(float* __restrict__ vol, const float* __restrict__ maskf,
                                   int ZN, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    vol[i] *= maskf[i / ZN];
}

__global__ void from_bf16_kernel(float* __restrict__ dst, const uint16_t* __restrict__ src,
                                 size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = ld(src, i);
}

// out = (in != 0) ? 1/in : 1  -- the guarded divisions of os_sart.m:30-31,38-39.
__global__ void inv_guard_kernel(const float* __restrict__ in, float* __restrict__ out,
                                 size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float a = in[i];
    out[i] = (a != 0.f) ? 1.0f / a : 1.0f;
  }
}

__global__ void axpy_kernel(float* __restrict__ y, const float* __restrict__ x, double a,
                            size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = (float)((double)y[i] + a * (double)x[i]);
}

__global__ void xpby_kernel(float* __restrict__ p, const float* __restrict__ r, double b,
                            size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = (float)((double)r[i] + b * (double)p[i]);
}

__global__ void sub_kernel(float* __restrict__ out, const float* __restrict__ a,
                           const float* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = a[i] - b[i];
}

// FP64 dot product: per-block partial sums (fixed order), then one block reduces.
__global__ __launch_bounds__(kBlock) void dot_partial_kernel(const float* __restrict__ a,
                                                             const float* __restrict__ b,
                                                             size_t n,
                                                             double* __restrict__ partials) {
  double s = 0.0;
  for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    s += (double)a[i] * (double)b[i];
  block_sum2(s, 0.0, partials + 2 * blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void reduce_pairs_kernel(const double* __restrict__ partials,
                                                              int nparts,
                                                              double* __restrict__ result) {
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kBlock) {
    a += partials[2 * i];
    b += partials[2 * i + 1];
  }
  block_sum2(a, b, result);
}

// One block per scan: items[blockIdx.x].out[0..1] = sums of its partial pairs.
__global__ __launch_bounds__(kBlock) void reduce_batch_kernel(const ReduceItem* __restrict__ items) {
  const ReduceItem it = items[blockIdx.x];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < it.n; i += kBlock) {
    a += it.partials[2 * i];
    b += it.partials[2 * i + 1];
  }
  block_sum2(a, b, it.out);
}

int grid_1d(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CTR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return CTR_OK;
}

}  // namespace

int fp_partial_blocks(const Plan& P, int vc) {
  return ((P.d.DNU * P.d.DNV + kBlock - 1) / kBlock) * vc;
}

template <typename TV, typename TO>
FpArgs<TV, TO> fp_args(const Plan& P, const TV* vol, TO* out, int vb, int vs, const float* b,
                       const float* invRow, double* partials) {
  return FpArgs<TV, TO>{vol,     out,    P.views, P.bnd, P.bnd64, P.cols,  P.Bz,
                        P.h2,    P.invdB, P.maskf, P.mext, P.d,   vb,      vs,
                        b,       invRow, partials};
}

template <typename T>
BpArgs<T> bp_args(const Plan& P, const T* prj, float* vol, int vb, int vs, int vc,
                  const float* invCol, uint16_t* shadow) {
  return BpArgs<T>{prj, vol, P.views, P.src64, P.mask, P.d, P.bc, P.bc64, vb, vs, vc,
                   invCol, shadow};
}

int fp_blocks_per_view(const Plan& P, bool colw) {
  return colw ? (P.d.DNU + kWaves - 1) / kWaves : (P.d.DNU * P.d.DNV + kBlock - 1) / kBlock;
}

template <typename TV, typename TO>
int launch_fp(const Plan& P, const TV* vol, TO* out, int vb, int vs, int vc, hipStream_t st,
              const float* b, const float* invRow, double* partials) {
  if (vc <= 0) return CTR_OK;
  // with DNV == 64 both mappings have DNU * 64 / 256 blocks per view
  dim3 grid((P.d.DNU * P.d.DNV + kBlock - 1) / kBlock, vc);
  const bool colw = P.d.DNV == 64;
  const FpArgs<TV, TO> A = fp_args<TV, TO>(P, vol, out, vb, vs, b, invRow, partials);
  if (b) {
    if (colw) hipLaunchKernelGGL((fp_kernel<true, true, TV, TO>), grid, dim3(kBlock), 0, st, A);
    else hipLaunchKernelGGL((fp_kernel<true, false, TV, TO>), grid, dim3(kBlock), 0, st, A);
  } else {
    if (colw) hipLaunchKernelGGL((fp_kernel<false, true, TV, TO>), grid, dim3(kBlock), 0, st, A);
    else hipLaunchKernelGGL((fp_kernel<false, false, TV, TO>), grid, dim3(kBlock), 0, st, A);
  }
  return check_launch("fp_kernel");
}

int launch_fp_ones(const Plan& P, float* out, int vb, int vs, int vc, hipStream_t st) {
  if (vc <= 0) return CTR_OK;
  dim3 grid((P.d.DNU * P.d.DNV + kBlock - 1) / kBlock, vc);
  const FpArgs<float, float> A =
      fp_args<float, float>(P, nullptr, out, vb, vs, nullptr, nullptr, nullptr);
  if (P.d.DNV == 64) hipLaunchKernelGGL((fp_ones_kernel<true>), grid, dim3(kBlock), 0, st, A);
  else hipLaunchKernelGGL((fp_ones_kernel<false>), grid, dim3(kBlock), 0, st, A);
  return check_launch("fp_ones_kernel");
}

int launch_bp_ones(const Plan& P, float* vol, int vb, int vs, int vc, hipStream_t st) {
  if (P.d.ZN > kWinMaxZ) return fail(CTR_ERR_ARG, "bp of ones: columns over the LDS limit");
  const BpArgs<float> A = bp_args<float>(P, nullptr, vol, vb, vs, vc, nullptr, nullptr);
  const dim3 wgrid((P.d.XN * P.d.YN + kWaves - 1) / kWaves);
  hipLaunchKernelGGL(bp_win_ones_kernel, wgrid, dim3(kBlock),
                     (size_t)kWaves * (kOnesRecF + P.d.ZN) * sizeof(float), st, A);
  return check_launch("bp_win_ones_kernel");
}

template <typename T>
int launch_bp(const Plan& P, const T* prj, float* vol, int vb, int vs, int vc, int kind,
              hipStream_t st, const float* invCol, int accumulate, uint16_t* shadow) {
  const int zg = std::min(4, (P.d.ZN + 63) / 64);
  dim3 rgrid((P.d.XN * P.d.YN + kWaves - 1) / kWaves, (P.d.ZN + 64 * zg - 1) / (64 * zg));
  if (kind == CTR_BP_MATCHED) {
    if constexpr (!std::is_same<T, float>::value) {
      return fail(CTR_ERR_ARG, "the matched back projector reads the FP32 scaled sinogram");
    } else {
      // windowed kernel while its row table and column sums fit 64 KiB of LDS
      const size_t lds =
          ((size_t)((P.d.DNV + 1 + 3) & ~3) + (size_t)kWaves * P.d.ZN) * sizeof(float);
      if (P.d.ZN <= kWinMaxZ && lds <= 65536) {
        const dim3 wgrid((P.d.XN * P.d.YN + kWaves - 1) / kWaves);
        if (accumulate)
          hipLaunchKernelGGL((bp_matched_win_kernel<true>), wgrid, dim3(kBlock), lds, st, prj,
                             vol, P.views, P.bnd64, P.cols, P.Bz, P.mask, P.d, P.bc, vb, vs, vc,
                             invCol, shadow);
        else
          hipLaunchKernelGGL((bp_matched_win_kernel<false>), wgrid, dim3(kBlock), lds, st, prj,
                             vol, P.views, P.bnd64, P.cols, P.Bz, P.mask, P.d, P.bc, vb, vs, vc,
                             nullptr, nullptr);
        return check_launch("bp_matched_win_kernel");
      }
#define CTR_BPM(UPD, ZGV, IC, SH)                                                                \
  hipLaunchKernelGGL((bp_matched_kernel<UPD, ZGV>), rgrid, dim3(kBlock), 0, st, prj, vol, P.views, \
                     P.bnd64, P.cols, P.Bz, P.mask, P.d, P.bc, vb, vs, vc, IC, SH)
      if (accumulate) {
        switch (zg) {
          case 1: CTR_BPM(true, 1, invCol, shadow); break;
          case 2: CTR_BPM(true, 2, invCol, shadow); break;
          case 3: CTR_BPM(true, 3, invCol, shadow); break;
          default: CTR_BPM(true, 4, invCol, shadow); break;
        }
      } else {
        switch (zg) {
          case 1: CTR_BPM(false, 1, nullptr, nullptr); break;
          case 2: CTR_BPM(false, 2, nullptr, nullptr); break;
          case 3: CTR_BPM(false, 3, nullptr, nullptr); break;
          default: CTR_BPM(false, 4, nullptr, nullptr); break;
        }
      }
#undef CTR_BPM
      return check_launch("bp_matched_kernel");
    }
  }
  const BpArgs<T> A = bp_args<T>(P, prj, vol, vb, vs, vc, accumulate ? invCol : nullptr,
                                 accumulate ? shadow : nullptr);
  if (P.d.ZN <= kWinMaxZ) {
    const dim3 wgrid((P.d.XN * P.d.YN + kWaves - 1) / kWaves);
    const size_t stg = P.d.DNV <= kStageRows ? kStageF + kLaccPad : 0;  // staged views
    const size_t lds = (size_t)kWaves * (P.d.ZN + stg) * sizeof(float);
    if (accumulate)
      hipLaunchKernelGGL((bp_win_kernel<true, T>), wgrid, dim3(kBlock), lds, st, A);
    else
      hipLaunchKernelGGL((bp_win_kernel<false, T>), wgrid, dim3(kBlock), lds, st, A);
    return check_launch("bp_win_kernel");
  }
#define CTR_BPREF(UPD, ZGV) \
  hipLaunchKernelGGL((bp_ref_kernel<UPD, ZGV, T>), rgrid, dim3(kBlock), 0, st, A)
  if (accumulate) {
    switch (zg) {
      case 1: CTR_BPREF(true, 1); break;
      case 2: CTR_BPREF(true, 2); break;
      case 3: CTR_BPREF(true, 3); break;
      default: CTR_BPREF(true, 4); break;
    }
  } else {
    switch (zg) {
      case 1: CTR_BPREF(false, 1); break;
      case 2: CTR_BPREF(false, 2); break;
      case 3: CTR_BPREF(false, 3); break;
      default: CTR_BPREF(false, 4); break;
    }
  }
#undef CTR_BPREF
  return check_launch("bp_ref_kernel");
}

// grid <= 0: one workgroup per item (hardware scheduling, per-scan XCD map)
template <typename T>
int launch_fp_batch(const FpBatchItem<T>* items, int nscan, int total, bool colw, int grid,
                    int strided, hipStream_t st) {
  if (total <= 0) return CTR_OK;
  const int sched = grid <= 0 ? 2 : (strided ? 1 : 0);
  grid = grid <= 0 ? total : std::max(1, std::min(grid, total));
  if (sched == 2 && colw)
    hipLaunchKernelGGL((fp_batch1_kernel<true, true, T>), dim3(grid), dim3(kBlock), 0, st,
                       items, nscan, total);
  else if (sched == 2)
    hipLaunchKernelGGL((fp_batch1_kernel<true, false, T>), dim3(grid), dim3(kBlock), 0, st,
                       items, nscan, total);
  else if (colw)
    hipLaunchKernelGGL((fp_batch_kernel<true, true, T>), dim3(grid), dim3(kBlock), 0, st, items,
                       nscan, total, sched);
  else
    hipLaunchKernelGGL((fp_batch_kernel<true, false, T>), dim3(grid), dim3(kBlock), 0, st, items,
                       nscan, total, sched);
  return check_launch("fp_batch_kernel");
}

template <typename T>
int launch_bp_batch(const BpBatchItem<T>* items, int nscan, int total, int zg, int grid,
                    int strided, hipStream_t st) {
  if (total <= 0) return CTR_OK;
  const int sched = grid <= 0 ? 2 : (strided ? 1 : 0);
  grid = grid <= 0 ? total : std::max(1, std::min(grid, total));
#define CTR_BPB(ZGV)                                                                       \
  hipLaunchKernelGGL((bp_batch_kernel<true, ZGV, T>), dim3(grid), dim3(kBlock), 0, st, items, \
                     nscan, total, sched)
  switch (zg) {
    case 1: CTR_BPB(1); break;
    case 2: CTR_BPB(2); break;
    case 3: CTR_BPB(3); break;
    case 4: CTR_BPB(4); break;
    default: return fail(CTR_ERR_ARG, "bp_batch: z groups must be 1..4");
  }
#undef CTR_BPB
  return check_launch("bp_batch_kernel");
}

template <typename T>
int launch_bp_batch_win(const BpBatchItem<T>* items, int nscan, int total, int zmax,
                        int maxDNV, int grid, int strided, hipStream_t st) {
  if (total <= 0) return CTR_OK;
  if (zmax < 1 || zmax > kWinMaxZ) return fail(CTR_ERR_ARG, "bp_batch_win: column height");
  const int sched = grid <= 0 ? 2 : (strided ? 1 : 0);
  grid = grid <= 0 ? total : std::max(1, std::min(grid, total));
  const int staged = maxDNV <= kStageRows ? 1 : 0;
  const size_t lds =
      (size_t)kWaves * (zmax + (staged ? kStageF + kLaccPad : 0)) * sizeof(float);
  if (sched == 2)
    hipLaunchKernelGGL((bp_batch_win1_kernel<T>), dim3(grid), dim3(kBlock), lds, st, items,
                       nscan, total, zmax, staged);
  else
    hipLaunchKernelGGL((bp_batch_win_kernel<T>), dim3(grid), dim3(kBlock), lds, st, items,
                       nscan, total, sched, zmax, staged);
  return check_launch("bp_batch_win_kernel");
}

int launch_reduce_batch(const ReduceItem* items, int nscan, hipStream_t st) {
  hipLaunchKernelGGL(reduce_batch_kernel, dim3(nscan), dim3(kBlock), 0, st, items);
  return check_launch("reduce_batch_kernel");
}

template <typename T>
int launch_scale_matched(const Plan& P, const T* in, float* out, int vb, int vs, int vc,
                         hipStream_t st) {
  if (vc <= 0) return CTR_OK;
  dim3 grid((P.d.DNU * P.d.DNV + kBlock - 1) / kBlock, vc);
  hipLaunchKernelGGL(scale_matched_kernel<T>, grid, dim3(kBlock), 0, st, in, out, P.cols, P.h2,
                     P.invdB, P.d, vb, vs);
  return check_launch("scale_matched_kernel");
}

template int launch_fp<float, float>(const Plan&, const float*, float*, int, int, int,
                                     hipStream_t, const float*, const float*, double*);
template int launch_fp<uint16_t, uint16_t>(const Plan&, const uint16_t*, uint16_t*, int, int,
                                           int, hipStream_t, const float*, const float*,
                                           double*);
template int launch_fp<float, uint16_t>(const Plan&, const float*, uint16_t*, int, int, int,
                                        hipStream_t, const float*, const float*, double*);
template int launch_bp<float>(const Plan&, const float*, float*, int, int, int, int, hipStream_t,
                              const float*, int, uint16_t*);
template int launch_bp<uint16_t>(const Plan&, const uint16_t*, float*, int, int, int, int,
                                 hipStream_t, const float*, int, uint16_t*);
template FpArgs<float> fp_args<float, float>(const Plan&, const float*, float*, int, int,
                                             const float*, const float*, double*);
template FpArgs<uint16_t> fp_args<uint16_t, uint16_t>(const Plan&, const uint16_t*, uint16_t*,
                                                      int, int, const float*, const float*,
                                                      double*);
template BpArgs<float> bp_args<float>(const Plan&, const float*, float*, int, int, int,
                                      const float*, uint16_t*);
template BpArgs<uint16_t> bp_args<uint16_t>(const Plan&, const uint16_t*, float*, int, int, int,
                                            const float*, uint16_t*);
template int launch_fp_batch<float>(const FpBatchItem<float>*, int, int, bool, int, int,
                                    hipStream_t);
template int launch_fp_batch<uint16_t>(const FpBatchItem<uint16_t>*, int, int, bool, int, int,
                                       hipStream_t);
template int launch_bp_batch<float>(const BpBatchItem<float>*, int, int, int, int, int,
                                    hipStream_t);
template int launch_bp_batch_win<float>(const BpBatchItem<float>*, int, int, int, int, int,
                                        int, hipStream_t);
template int launch_bp_batch_win<uint16_t>(const BpBatchItem<uint16_t>*, int, int, int, int,
                                           int, int, hipStream_t);
template int launch_bp_batch<uint16_t>(const BpBatchItem<uint16_t>*, int, int, int, int, int,
                                       hipStream_t);
template int launch_scale_matched<float>(const Plan&, const float*, float*, int, int, int,
                                         hipStream_t);
template int launch_scale_matched<uint16_t>(const Plan&, const uint16_t*, float*, int, int, int,
                                            hipStream_t);

int launch_to_bf16(uint16_t* dst, const float* src, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_1d(n)), dim3(256), 0, st, dst, src, n);
  return check_launch("to_bf16_kernel");
}

int launch_to_bf16_shadow(uint16_t* dst, const float* src, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(to_bf16_dither_kernel, dim3(grid_1d(n)), dim3(256), 0, st, dst, src, n);
  return check_launch("to_bf16_dither_kernel");
}

int launch_mask_volume(const Plan& P, float* vol, hipStream_t st) {
  const size_t n = (size_t)P.d.XN * P.d.YN * P.d.ZN;
  hipLaunchKernelGGL(mask_volume_kernel, dim3(grid_1d(n)), dim3(256), 0, st, vol, P.maskf,
                     P.d.ZN, n);
  return check_launch("mask_volume_kernel");
}

int launch_from_bf16(float* dst, const uint16_t* src, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(from_bf16_kernel, dim3(grid_1d(n)), dim3(256), 0, st, dst, src, n);
  return check_launch("from_bf16_kernel");
}

int launch_copy(float* dst, const float* src, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(copy_kernel, dim3(grid_1d(n)), dim3(256), 0, st, dst, src, n);
  return check_launch("copy_kernel");
}

int launch_fill(float* p, size_t n, float v, hipStream_t st) {
  hipLaunchKernelGGL(fill_kernel, dim3(grid_1d(n)), dim3(256), 0, st, p, n, v);
  return check_launch("fill_kernel");
}

int launch_inv_guard(const float* in, float* out, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(inv_guard_kernel, dim3(grid_1d(n)), dim3(256), 0, st, in, out, n);
  return check_launch("inv_guard_kernel");
}

int launch_axpy(float* y, const float* x, double a, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(axpy_kernel, dim3(grid_1d(n)), dim3(256), 0, st, y, x, a, n);
  return check_launch("axpy_kernel");
}

int launch_xpby(float* p, const float* r, double b, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(xpby_kernel, dim3(grid_1d(n)), dim3(256), 0, st, p, r, b, n);
  return check_launch("xpby_kernel");
}

int launch_sub(float* out, const float* a, const float* b, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(sub_kernel, dim3(grid_1d(n)), dim3(256), 0, st, out, a, b, n);
  return check_launch("sub_kernel");
}

// result[0] = sum(a*b) in FP64; partials holds 2*nparts doubles.
int launch_dot(const float* a, const float* b, size_t n, double* partials, int nparts,
               double* result, hipStream_t st) {
  hipLaunchKernelGGL(dot_partial_kernel, dim3(nparts), dim3(kBlock), 0, st, a, b, n, partials);
  int rc = check_launch("dot_partial_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(reduce_pairs_kernel, dim3(1), dim3(kBlock), 0, st, partials, nparts, result);
  return check_launch("reduce_pairs_kernel");
}

int launch_reduce_pairs(const double* partials, int nparts, double* result, hipStream_t st) {
  hipLaunchKernelGGL(reduce_pairs_kernel, dim3(1), dim3(kBlock), 0, st, partials, nparts, result);
  return check_launch("reduce_pairs_kernel");
}

}  // namespace ctr
