// gfx950 kernels of the reference's other projector models: the pseudo distance-driven
// pair (iterativeRecon/DD3_GPU.cu:673-834 forward, :1478-1528 back; prjMode 1 / 3
// forward, 1 / 2 back) and the z-line back projector (:1545-1647, prjMode 3 back).
//
// The reference samples linear-filtered, border-addressed textures (zero outside the
// array).  Here every sample is an explicit bilinear fetch of the array with exact FP32
// weights: taps outside the array get weight 0 and are never read out of bounds.
//
// Work mapping as in the distance-driven kernels: a wave64 owns 64 consecutive
// z-cells -- detector rows of one (view, column) in the forward projector, voxels of
// one (x, y) column in the back projector -- so the in-plane part of every sample is
// uniform over the wave and only z is per lane.  Uniform per-slice (forward) or
// per-view (back) quantities are evaluated 64 at a time, one slice / view per lane,
// and read back with readlane inside the loop over the ones that matter.
#include <algorithm>

#include "ctr_internal.hpp"
#include "device_common.hpp"

namespace ctr {
namespace {

__device__ __forceinline__ float rdl(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Bilinear taps of coordinate t (texel-index units, i.e. the texture coordinate minus
// 0.5) along an axis of n texels with zero outside: first texel i0 (clamped into the
// array for the read), weights w0 (texel floor(t)) and w1 (texel floor(t) + 1).
__device__ __forceinline__ void taps(float t, int n, int& i0, int& i1, float& w0, float& w1) {
  const float f = floorf(t);
  const int k = (int)f;
  const float a = t - f;
  w0 = (k >= 0 && k < n) ? 1.f - a : 0.f;
  w1 = (k + 1 >= 0 && k + 1 < n) ? a : 0.f;
  i0 = min(max(k, 0), n - 1);
  i1 = min(max(k + 1, 0), n - 1);
}

// ---------------------------------------------------------------------------
// Forward (DD3_GPU.cu:673-754).  Ray from the source to the centre of cell (u, v);
// on each slice ii of the marching axis (x when |cos| <= |sin|, else y) it crosses the
// slice's centre plane at lateral index li(ii) = A + ii B and z index
// zi(ii) = az + r(ii) zc_v, r(ii) = r0 + ii r1 being the fraction of the source ->
// cell vector; the masked volume is sampled there bilinearly in (lateral, z) and the
// sum is scaled by dx |e| / |e_along| (e = cell centre - source).
// A, B are evaluated in FP64 from the rotated source and cell (the lateral index must
// resolve well below a cell across wide volumes); r0, r1, |e_xy|^2 and dx / |e_along|
// come from the distance-driven column table.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void fp_pdd_kernel(
    const float* __restrict__ vol, float* __restrict__ out, const ViewP* __restrict__ views,
    const double2* __restrict__ src64, const double2* __restrict__ rot64,
    const double2* __restrict__ det64, const float4* __restrict__ cols,
    const float* __restrict__ zc, const float* __restrict__ h2, const float* __restrict__ maskf,
    Dims d, int vb, int vs, int nvc, int items) {
  const int hw = blockIdx.x;
  const int lb = xcd_logical_block(hw, gridDim.x);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int item = lb * kWaves + wid;
  if (item >= items) return;
  const int lane = threadIdx.x & 63;
  // item = (k * nvc + vc) * DNU + u: neighbouring waves take neighbouring columns
  const int u = item % d.DNU;
  const int kv = item / d.DNU;
  const int k = kv / nvc;
  const int vc = kv - k * nvc;
  const int p = vb + k * vs;
  const int v = vc * 64 + lane;
  const bool vok = v < d.DNV;
  const int vr = vok ? v : d.DNV - 1;

  const ViewP vp = views[p];
  const double2 cs = rot64[p], s = src64[p], dd = det64[u];
  const double Dx = dd.x * cs.x - dd.y * cs.y, Dy = dd.x * cs.y + dd.y * cs.x;
  const double ex = Dx - s.x, ey = Dy - s.y;
  const float4 ct = cols[(size_t)p * d.DNU + u];  // (r0, r1, |e_xy|^2, dx / |e_along|)
  double A, B;
  int NS, NL, sS, sL;
  if (vp.xmarch) {
    const double r0 = (-(double)d.ocx * d.dx - s.x) / ex, r1 = (double)d.dx / ex;
    A = (s.y + r0 * ey) / d.dx + d.ocy;
    B = r1 * ey / d.dx;
    NS = d.XN; NL = d.YN; sS = d.ZN; sL = d.XN * d.ZN;
  } else {
    const double r0 = (-(double)d.ocy * d.dx - s.y) / ey, r1 = (double)d.dx / ey;
    A = (s.x + r0 * ex) / d.dx + d.ocx;
    B = r1 * ex / d.dx;
    NS = d.YN; NL = d.XN; sS = d.XN * d.ZN; sL = d.ZN;
  }
  const int mS = vp.xmarch ? 1 : d.XN, mL = vp.xmarch ? d.XN : 1;
  const float zcv = zc[vr];
  // z extent of the wave's rays (zc increases with the row), for the slice test
  const int vlast = min(vc * 64 + 63, d.DNV - 1);
  const float zc_lo = zc[vc * 64], zc_hi = zc[vlast];
  const float ZN = (float)d.ZN;
  const Buf<float> V(vol, (uint32_t)(d.XN * d.YN * d.ZN));
  float acc = 0.f;
  for (int i0 = 0; i0 < NS; i0 += 64) {
    const int il = i0 + lane;
    int q_o0 = 0, q_o1 = 0;
    float q_w0 = 0.f, q_w1 = 0.f, q_r = 0.f;
    bool ok = false;
    if (il < NS) {
      const double li = fma((double)il, B, A);
      const float r = fmaf((float)il, ct.y, ct.x);
      const float za = fmaf(r, zc_lo, vp.az), zb = fmaf(r, zc_hi, vp.az);
      if (li > -1.0 && li < (double)NL && fmaxf(za, zb) > -1.f && fminf(za, zb) < ZN) {
        const double fl = floor(li);
        const int j = (int)fl;
        const float a = (float)(li - fl);
        const int j0 = max(j, 0), j1 = min(j + 1, NL - 1);
        q_w0 = (j >= 0) ? (1.f - a) * maskf[il * mS + j0 * mL] : 0.f;
        q_w1 = (j + 1 < NL) ? a * maskf[il * mS + j1 * mL] : 0.f;
        q_o0 = il * sS + j0 * sL;
        q_o1 = il * sS + j1 * sL;
        q_r = r;
        ok = (q_w0 != 0.f) || (q_w1 != 0.f);
      }
    }
    uint64_t live = __ballot(ok);
    while (live) {
      const int l = __builtin_ctzll(live);
      live &= live - 1;
      const float w0 = rdl(q_w0, l), w1 = rdl(q_w1, l), r = rdl(q_r, l);
      const int o0 = __builtin_amdgcn_readlane(q_o0, l), o1 = __builtin_amdgcn_readlane(q_o1, l);
      int k0, k1;
      float wz0, wz1;
      taps(fmaf(r, zcv, vp.az), d.ZN, k0, k1, wz0, wz1);
      const float a0 = V[o0 + k0], a1 = V[o0 + k1], b0 = V[o1 + k0], b1 = V[o1 + k1];
      acc = fmaf(w0, fmaf(wz0, a0, wz1 * a1), acc);
      acc = fmaf(w1, fmaf(wz0, b0, wz1 * b1), acc);
    }
  }
  if (vok)
    out[((size_t)k * d.DNU + u) * d.DNV + v] =
        acc * ct.w * __builtin_amdgcn_sqrtf(ct.z + h2[v]);
}

// ---------------------------------------------------------------------------
// Back (DD3_GPU.cu:1478-1528).  For voxel P and view p with source s, d = P - s:
//   iu = asin((s_y d_x - s_x d_y) / (|d_xy| |s_xy|)) / dbeta + cU     (FP64, uniform)
//   iv = d_z (S2D / ddv) / |d_xy| + cV                               (FP32, per lane)
//   w  = dx |d| / |d_x| if |s_x| >= |s_y| else dx |d| / |d_y|
// and the voxel gets the sum over views of w times the bilinear sample of view p at
// (iu, iv).  Voxels whose mask byte is not 1 are 0.
// ---------------------------------------------------------------------------
template <int ZG>
__global__ __launch_bounds__(kBlock) void bp_pdd_kernel(
    const float* __restrict__ prj, float* __restrict__ vol, const ViewP* __restrict__ views,
    const double2* __restrict__ src64, const uint8_t* __restrict__ mask, Dims d, BpConsts c,
    BpConsts64 c64, int vb, int vs, int vc) {
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (col >= d.XN * d.YN) return;
  const int zb0 = blockIdx.y * (64 * ZG);
  const int zb = zb0 + lane;
  const int y = col / d.XN, x = col - y * d.XN;
  float* vcol = vol + (size_t)col * d.ZN;
  if (mask[col] != 1) {
#pragma unroll
    for (int g = 0; g < ZG; ++g)
      if (zb + 64 * g < d.ZN) vcol[zb + 64 * g] = 0.f;
    return;
  }
  const double Px = ((double)x - c64.ocx) * c64.dx;
  const double Py = ((double)y - c64.ocy) * c64.dx;
  const float Pxf = (float)Px, Pyf = (float)Py;
  const float Pz0 = ((float)zb - d.ocz) * d.dz;
  const float dzg = 64.f * d.dz;
  const float Zlo = ((float)zb0 - d.ocz) * d.dz;
  const float Zhi = ((float)min(zb0 + 64 * ZG, d.ZN) - 1.f - d.ocz) * d.dz;
  const float DNV = (float)d.DNV;
  const uint32_t frame = (uint32_t)(d.DNU * d.DNV);
  float acc[ZG];
#pragma unroll
  for (int g = 0; g < ZG; ++g) acc[g] = 0.f;
  for (int k0 = 0; k0 < vc; k0 += 64) {
    const int kl = k0 + lane;
    float q_rlS = 0.f, q_sz = 0.f, q_l2 = 0.f, q_cw = 0.f, q_wu0 = 0.f, q_wu1 = 0.f;
    int q_u0 = 0, q_u1 = 0;
    bool vis = false;
    if (kl < vc) {
      const int p = vb + kl * vs;
      const ViewP vp = views[p];
      const double2 s64 = src64[p];
      const double dX = Px - s64.x, dY = Py - s64.y;
      const double l2 = dX * dX + dY * dY;
      const double rl = rsqrt(l2);
      const double iu = asin_fan((s64.y * Px - s64.x * Py) * c64.inv_sid * rl) * c64.ib + c64.cU;
      const float dXf = Pxf - vp.sx, dYf = Pyf - vp.sy;
      const float l2f = dXf * dXf + dYf * dYf;
      const float rlS = __builtin_amdgcn_rsqf(l2f) * c.S;
      const float ta = fmaf(Zlo - vp.sz, rlS, c.cV), tb = fmaf(Zhi - vp.sz, rlS, c.cV);
      vis = iu > -1.0 && iu < (double)d.DNU && fmaxf(ta, tb) > -1.f && fminf(ta, tb) < DNV;
      if (vis) {
        const double fu = floor(iu);
        const int u = (int)fu;
        const float a = (float)(iu - fu);
        q_wu0 = (u >= 0) ? 1.f - a : 0.f;
        q_wu1 = (u + 1 < d.DNU) ? a : 0.f;
        q_u0 = max(u, 0);
        q_u1 = min(u + 1, d.DNU - 1);
        const double w = (fabs(s64.x) >= fabs(s64.y)) ? dX : dY;
        q_cw = c.dx * __builtin_amdgcn_rcpf(fabsf((float)w));
        q_rlS = rlS;
        q_sz = vp.sz;
        q_l2 = l2f;
      }
    }
    uint64_t live = __ballot(vis);
    while (live) {
      const int j = __builtin_ctzll(live);
      live &= live - 1;
      const int k = k0 + j;
      const float rlS = rdl(q_rlS, j), vsz = rdl(q_sz, j), l2f = rdl(q_l2, j);
      const float cw = rdl(q_cw, j), wu0 = rdl(q_wu0, j), wu1 = rdl(q_wu1, j);
      const int u0 = __builtin_amdgcn_readlane(q_u0, j), u1 = __builtin_amdgcn_readlane(q_u1, j);
      const Buf<float> pp(prj + (size_t)k * frame, frame);
      const int o0 = u0 * d.DNV, o1 = u1 * d.DNV;
#pragma unroll
      for (int g = 0; g < ZG; ++g) {
        const float dZ = fmaf((float)g, dzg, Pz0) - vsz;
        const float iv = fmaf(dZ, rlS, c.cV);
        const bool ok = zb + 64 * g < d.ZN && iv > -1.f && iv < DNV;
        if (!__any(ok)) continue;
        if (!ok) continue;
        int v0, v1;
        float wv0, wv1;
        taps(iv, d.DNV, v0, v1, wv0, wv1);
        const float s = wu0 * fmaf(wv0, pp[o0 + v0], wv1 * pp[o0 + v1]) +
                        wu1 * fmaf(wv0, pp[o1 + v0], wv1 * pp[o1 + v1]);
        acc[g] = fmaf(cw * __builtin_amdgcn_sqrtf(fmaf(dZ, dZ, l2f)), s, acc[g]);
      }
    }
  }
#pragma unroll
  for (int g = 0; g < ZG; ++g)
    if (zb + 64 * g < d.ZN) vcol[zb + 64 * g] = acc[g];
}

// ---------------------------------------------------------------------------
// z-line back projector (DD3_GPU.cu:1545-1647), prjMode 3.
//
// The reference accumulates, per voxel column and block of 16 slices, 17 face terms
//   T_p(j) = cosVal_p w_p(j) [SAT_p(tD_j, tR) - SAT_p(tD_j, tL)]
// over the views (face j: the lower face of slice j; w_p(j) = hypot(d_z, |d_xy|)/|d_xy|
// at the centre of slice j; [tL, tR]: fan-angle range of the column's extent across the
// rays), and writes slice k = S(k+1) - S(k), S(j) = sum_p T_p(j), for the slices of the
// whole blocks (k < 16 * floor(ZN / 16)); the others stay 0.
//
// Here a wave owns one column and its lanes the faces; S is accumulated in FP64 (the
// slice values are differences of sums that carry whole detector columns) and the
// bracket is the exact integral of the piecewise-constant view over the rows below tD
// and the columns of [tL, tR]: per column u, C_u(m) + frac * prj[u][m], C_u the
// exclusive FP64 prefix over the rows (zl_prefix_kernel).  The upper face of a wave's
// last slice is one more face, taken by lane 0 of an extra group.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void zl_prefix_kernel(const float* __restrict__ prj,
                                                           double* __restrict__ cpre, int rows,
                                                           int DNV) {
  const int row = blockIdx.x * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (row >= rows) return;
  const int lane = threadIdx.x & 63;
  const size_t base = (size_t)row * DNV;
  double carry = 0.0;
  for (int c0 = 0; c0 < DNV; c0 += 64) {
    const int v = c0 + lane;
    const double x = v < DNV ? (double)prj[base + v] : 0.0;
    double inc = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const double t = __shfl_up(inc, off, 64);
      if (lane >= off) inc += t;
    }
    if (v < DNV) cpre[base + v] = carry + inc - x;
    carry += __shfl(inc, 63, 64);
  }
}

template <int ZG>
__global__ __launch_bounds__(kBlock) void bp_zline_kernel(
    const float* __restrict__ prj, const double* __restrict__ cpre, float* __restrict__ vol,
    const ViewP* __restrict__ views, const double2* __restrict__ src64,
    const double* __restrict__ sz64, const uint8_t* __restrict__ mask, Dims d, BpConsts64 c64,
    int vb, int vs, int vc, int NZ) {
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (col >= d.XN * d.YN) return;
  const int zb0 = blockIdx.y * (64 * ZG);
  const int y = col / d.XN, x = col - y * d.XN;
  float* vcol = vol + (size_t)col * d.ZN;
  if (mask[col] != 1 || zb0 >= NZ) {
#pragma unroll
    for (int g = 0; g < ZG; ++g)
      if (zb0 + 64 * g + lane < d.ZN) vcol[zb0 + 64 * g + lane] = 0.f;
    return;
  }
  const double Px = ((double)x - c64.ocx) * c64.dx;
  const double Py = ((double)y - c64.ocy) * c64.dx;
  const double hdx = 0.5 * c64.dx, dz = d.dz, ocz = c64.ocz, cV = c64.cV;
  // faces of the wave: zb0 .. ftop (ftop <= NZ)
  const int ftop = min(zb0 + 64 * ZG, NZ);
  const double DNVd = (double)d.DNV;
  double S[ZG + 1];
#pragma unroll
  for (int g = 0; g <= ZG; ++g) S[g] = 0.0;
  for (int k0 = 0; k0 < vc; k0 += 64) {
    const int kl = k0 + lane;
    double q_s2v = 0.0, q_sz = 0.0, q_ang = 0.0, q_cw = 0.0, q_wu0 = 0.0, q_wul = 0.0;
    int q_u0 = 0, q_nu = 0;
    bool vis = false;
    if (kl < vc) {
      const int p = vb + kl * vs;
      const ViewP vp = views[p];
      const double2 s = src64[p];
      const double dX = Px - s.x, dY = Py - s.y;
      const double s2v = sqrt(dX * dX + dY * dY);
      double lx, ly, rx, ry, cosv;
      if (vp.xmarch) {
        lx = dX; ly = dY - hdx; rx = dX; ry = dY + hdx;
        cosv = c64.dx * s2v / dX;
      } else {
        lx = dX + hdx; ly = dY; rx = dX - hdx; ry = dY;
        cosv = c64.dx * s2v / dY;
      }
      const double tL =
          asin_fan((s.y * lx - s.x * ly) * rsqrt(lx * lx + ly * ly) * c64.inv_sid) * c64.ib +
          c64.cU;
      const double tR =
          asin_fan((s.y * rx - s.x * ry) * rsqrt(rx * rx + ry * ry) * c64.inv_sid) * c64.ib +
          c64.cU;
      const double lo = fmin(fmax(fmin(tL, tR), -0.5), d.DNU - 0.5);
      const double hi = fmin(fmax(fmax(tL, tR), -0.5), d.DNU - 0.5);
      const double ang = c64.S / s2v;
      const double sz = sz64[p];
      // the rows below the wave's highest face must not all lie below the detector
      const double ta = (((double)zb0 - ocz) * dz - sz - 0.5 * dz) * ang + cV;
      const double tb = (((double)ftop - ocz) * dz - sz - 0.5 * dz) * ang + cV;
      vis = lo < hi && fmax(ta, tb) > -0.5;
      if (vis) {
        const int u0 = (int)floor(lo + 0.5);
        const int nu = min((int)floor(hi + 0.5), d.DNU - 1) - u0 + 1;
        q_wu0 = (nu == 1) ? hi - lo : ((double)u0 + 0.5) - lo;
        q_wul = hi - ((double)(u0 + nu - 1) - 0.5);
        q_u0 = u0;
        q_nu = nu;
        q_cw = (tR >= tL) ? cosv : -cosv;
        q_s2v = s2v;
        q_sz = sz;
        q_ang = ang;
      }
    }
    uint64_t live = __ballot(vis);
    while (live) {
      const int j = __builtin_ctzll(live);
      live &= live - 1;
      const int k = k0 + j;
      const double s2v = __shfl(q_s2v, j, 64), sz = __shfl(q_sz, j, 64);
      const double ang = __shfl(q_ang, j, 64), cw = __shfl(q_cw, j, 64);
      const double wu0 = __shfl(q_wu0, j, 64), wul = __shfl(q_wul, j, 64);
      const int u0 = __builtin_amdgcn_readlane(q_u0, j), nu = __builtin_amdgcn_readlane(q_nu, j);
      const float* pk = prj + (size_t)k * d.DNU * d.DNV;
      const double* ck = cpre + (size_t)k * d.DNU * d.DNV;
#pragma unroll
      for (int g = 0; g <= ZG; ++g) {
        const int f = zb0 + 64 * g + lane;
        const bool fok = (g < ZG) ? f <= NZ : (lane == 0 && f <= NZ);
        if (!__any(fok)) continue;
        if (!fok) continue;
        const double dirZ = ((double)f - ocz) * dz - sz;
        const double tD = (dirZ - 0.5 * dz) * ang + cV;
        if (!(tD > -0.5)) continue;
        const double tt = fmin(tD + 0.5, DNVd);
        const int m = min((int)tt, d.DNV - 1);
        const double fr = tt - (double)m;
        double G = 0.0;
        for (int iu = 0; iu < nu; ++iu) {
          const double wu = (iu == 0) ? wu0 : (iu == nu - 1 ? wul : 1.0);
          const size_t o = (size_t)(u0 + iu) * d.DNV + m;
          G = fma(wu, fma(fr, (double)pk[o], ck[o]), G);
        }
        S[g] = fma(cw * sqrt(fma(dirZ, dirZ, s2v * s2v)) / s2v, G, S[g]);
      }
    }
  }
#pragma unroll
  for (int g = 0; g < ZG; ++g) {
    const int z = zb0 + 64 * g + lane;
    double up = __shfl_down(S[g], 1, 64);
    const double first_next = __shfl(S[g + 1], 0, 64);
    if (lane == 63) up = first_next;
    if (z < d.ZN) vcol[z] = (z < NZ) ? (float)(up - S[g]) : 0.f;
  }
}

int check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(CTR_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return CTR_OK;
}

}  // namespace

int launch_fp_pdd(const Plan& P, const float* vol, float* out, int vb, int vs, int vc,
                  hipStream_t st) {
  const Dims& d = P.d;
  if (vc <= 0) return CTR_OK;
  const int nvc = (d.DNV + 63) / 64;
  const long long items = (long long)vc * nvc * d.DNU;
  if (items > (long long)INT32_MAX - kWaves)
    return fail(CTR_ERR_UNSUPPORTED, "pseudo-DD forward: too many rays in one launch");
  const int grid = (int)((items + kWaves - 1) / kWaves);
  hipLaunchKernelGGL(fp_pdd_kernel, dim3(grid), dim3(kBlock), 0, st, vol, out, P.views,
                     P.src64, P.rot64, P.det64, P.cols, P.zc, P.h2, P.maskf, d, vb, vs, nvc,
                     (int)items);
  return check("fp_pdd_kernel");
}

int launch_bp_pdd(const Plan& P, const float* prj, float* vol, int vb, int vs, int vc,
                  hipStream_t st) {
  const int zg = std::min(4, (P.d.ZN + 63) / 64);
  dim3 grid((P.d.XN * P.d.YN + kWaves - 1) / kWaves, (P.d.ZN + 64 * zg - 1) / (64 * zg));
#define CTR_BPP(ZGV)                                                                     \
  hipLaunchKernelGGL(bp_pdd_kernel<ZGV>, grid, dim3(kBlock), 0, st, prj, vol, P.views, \
                     P.src64, P.mask, P.d, P.bc, P.bc64, vb, vs, vc)
  switch (zg) {
    case 1: CTR_BPP(1); break;
    case 2: CTR_BPP(2); break;
    case 3: CTR_BPP(3); break;
    default: CTR_BPP(4); break;
  }
#undef CTR_BPP
  return check("bp_pdd_kernel");
}

size_t zline_prefix_doubles(const Dims& d, int vc) { return (size_t)vc * d.DNU * d.DNV; }

int launch_bp_zline(const Plan& P, const float* prj, double* cpre, float* vol, int vb, int vs,
                    int vc, hipStream_t st) {
  const Dims& d = P.d;
  const int rows = vc * d.DNU;
  if (rows > 0) {
    hipLaunchKernelGGL(zl_prefix_kernel, dim3((rows + kWaves - 1) / kWaves), dim3(kBlock), 0, st,
                       prj, cpre, rows, d.DNV);
    int rc = check("zl_prefix_kernel");
    if (rc) return rc;
  }
  const int NZ = 16 * (d.ZN / 16);
  const int zg = std::min(4, (d.ZN + 63) / 64);
  dim3 grid((d.XN * d.YN + kWaves - 1) / kWaves, (d.ZN + 64 * zg - 1) / (64 * zg));
#define CTR_BPZ(ZGV)                                                                          \
  hipLaunchKernelGGL(bp_zline_kernel<ZGV>, grid, dim3(kBlock), 0, st, prj, cpre, vol, P.views, \
                     P.src64, P.sz64, P.mask, d, P.bc64, vb, vs, vc, NZ)
  switch (zg) {
    case 1: CTR_BPZ(1); break;
    case 2: CTR_BPZ(2); break;
    case 3: CTR_BPZ(3); break;
    default: CTR_BPZ(4); break;
  }
#undef CTR_BPZ
  return check("bp_zline_kernel");
}

}  // namespace ctr
