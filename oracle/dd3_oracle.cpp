// DD3 ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference's projector models (the mode-0 distance-driven
// pair and the other prjMode values) and of the OS-SART / CG loops that drive mode 0.  Only the tests, the smoke check and
// the CPU-baseline leg of the benchmark load this file's library; the product
// (ctrecon_amd) never links, imports or calls it.
//
// Each routine is templated on the arithmetic type T:
//   T = double  -> the FP64 oracle that GPU results are checked against;
//   T = float   -> the FP32 CPU baseline (OpenMP over the host cores).
//
// Followed reference code (LittleRabbitRun/CTRecon):
//   geometry / boundaries  iterativeRecon/utilities.cuh:257-265 (regularizeAngle),
//                          :291-312 (ConstantForBackProjection), :340-361 (DD3Boundaries)
//   volume SAT             iterativeRecon/DD3_GPU.cu:33-50, 101-197
//   forward (mode 0)       iterativeRecon/DD3_GPU.cu:844-990, host 992-1126
//   detector parameters    iterativeRecon/DD3_GPU.cu:1189-1217 (calDetParas)
//   sinogram SAT           iterativeRecon/DD3_GPU.cu:1241-1323
//   back (mode 0)          iterativeRecon/DD3_GPU.cu:1397-1475, host 1651-1760
//   forward, pseudo-DD     iterativeRecon/DD3_GPU.cu:673-834 (prjMode 1 and 3)
//   forward, FP64 SAT      iterativeRecon/DD3_GPU.cu:255-660 (prjMode 2)
//   back, pseudo-DD        iterativeRecon/DD3_GPU.cu:1478-1528 (prjMode 1 and 2)
//   back, z-line           iterativeRecon/DD3_GPU.cu:1545-1647 (prjMode 3)
//   OS-SART                iterativeRecon/os_sart.m:13-44
//   CG                     iterativeRecon/conjgrad.m:4-27
//
// Texture fetches (cudaFilterModeLinear, cudaAddressModeClamp, unnormalised
// coordinates) are restated with exact bilinear weights: sample at (a-0.5,b-0.5),
// clamp each texel index to the array.  NVIDIA's fixed-point filter weights and
// --use_fast_math intrinsics are not reproduced (they are not pinned by anything
// in the reference tree).
//
// Besides the SAT transcription there is an independent "direct" restatement of
// both projectors (exact rectangle overlap, no integral image) and the exact
// transpose of the direct forward projector (the matched adjoint).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kTwoPi = 6.283185307179586;  // utilities.cuh:75

template <typename T>
struct Geo {
  T x0, y0, z0;
  int DNU, DNV;
  std::vector<T> xds, yds, zds, bxds, byds, bzds;
  T ocx, ocy, ocz;
  std::vector<T> cs, sn, zsh;  // per view: cos, sin of regularised angle, zshift
  int PN, XN, YN, ZN;
  T dx, dz;
  const uint8_t* mask;
};

// utilities.cuh:340-361
template <typename T>
void boundaries(int nb, const T* c, T* b) {
  if (nb >= 3) {
    b[0] = T(1.5) * c[0] - T(0.5) * c[1];
    for (int i = 1; i <= nb - 2; ++i) b[i] = T(0.5) * c[i - 1] + T(0.5) * c[i];
    b[nb - 1] = T(1.5) * c[nb - 2] - T(0.5) * c[nb - 3];
  } else {
    b[0] = c[0] - T(0.5);
    b[1] = c[0] + T(0.5);
  }
}

// utilities.cuh:257-265
template <typename T>
T regularize(T a) {
  while (a >= T(kTwoPi)) a -= T(kTwoPi);
  while (a < T(0)) a += T(kTwoPi);
  return a;
}

template <typename T>
Geo<T> make_geo(float x0, float y0, float z0, int DNU, int DNV, const float* xds,
                const float* yds, const float* zds, float cx, float cy, float cz,
                const float* angs, const float* zpos, int PN, int XN, int YN, int ZN,
                float dx, float dz, const uint8_t* mask) {
  Geo<T> g;
  g.x0 = x0; g.y0 = y0; g.z0 = z0;
  g.DNU = DNU; g.DNV = DNV;
  g.xds.assign(xds, xds + DNU);
  g.yds.assign(yds, yds + DNU);
  g.zds.assign(zds, zds + DNV);
  g.bxds.resize(DNU + 1); g.byds.resize(DNU + 1); g.bzds.resize(DNV + 1);
  boundaries<T>(DNU + 1, g.xds.data(), g.bxds.data());
  boundaries<T>(DNU + 1, g.yds.data(), g.byds.data());
  boundaries<T>(DNV + 1, g.zds.data(), g.bzds.data());
  g.dx = dx; g.dz = dz;
  // DD3_GPU.cu:1029-1031
  g.ocx = T((XN - 1.0) * 0.5) - T(cx) / g.dx;
  g.ocy = T((YN - 1.0) * 0.5) - T(cy) / g.dx;
  g.ocz = T((ZN - 1.0) * 0.5) - T(cz) / g.dz;
  g.cs.resize(PN); g.sn.resize(PN); g.zsh.resize(PN);
  for (int p = 0; p < PN; ++p) {
    T a = regularize<T>(T(angs[p]));
    g.cs[p] = std::cos(a);
    g.sn[p] = std::sin(a);
    g.zsh[p] = zpos[p];
  }
  g.PN = PN; g.XN = XN; g.YN = YN; g.ZN = ZN;
  g.mask = mask;
  return g;
}

// Bilinear fetch with clamp addressing on a 2-D array plane[h*W + w]; (a, b) are
// unnormalised texture coordinates along W and H.
template <typename T>
inline T tex2(const T* plane, int W, int H, T a, T b) {
  a -= T(0.5);
  b -= T(0.5);
  a = std::min(std::max(a, T(-1)), T(W));
  b = std::min(std::max(b, T(-1)), T(H));
  T fa = std::floor(a), fb = std::floor(b);
  int i = int(fa), j = int(fb);
  T al = a - fa, be = b - fb;
  int i0 = std::min(std::max(i, 0), W - 1), i1 = std::min(std::max(i + 1, 0), W - 1);
  int j0 = std::min(std::max(j, 0), H - 1), j1 = std::min(std::max(j + 1, 0), H - 1);
  return (T(1) - al) * (T(1) - be) * plane[(size_t)j0 * W + i0] +
         al * (T(1) - be) * plane[(size_t)j0 * W + i1] +
         (T(1) - al) * be * plane[(size_t)j1 * W + i0] + al * be * plane[(size_t)j1 * W + i1];
}

// DD3_GPU.cu:33-50 + 101-197: masked copy into two zero-padded arrays and 2-D
// prefix sums, ZXY[y][x'][z'] (per-y plane over (x, z)) and ZYX[x][y'][z'].
template <typename T>
void volume_sats(const Geo<T>& g, const float* vol, std::vector<T>& zxy, std::vector<T>& zyx) {
  const int XN = g.XN, YN = g.YN, ZN = g.ZN, W = ZN + 1;
  zxy.assign((size_t)YN * (XN + 1) * W, T(0));
  zyx.assign((size_t)XN * (YN + 1) * W, T(0));
  for (int y = 0; y < YN; ++y)
    for (int x = 0; x < XN; ++x) {
      T m = T(g.mask[y * XN + x]);  // DD3_GPU.cu:1004-1011 (vol *= mask)
      for (int z = 0; z < ZN; ++z) {
        T v = T(vol[((size_t)y * XN + x) * ZN + z]) * m;
        zxy[((size_t)y * (XN + 1) + x + 1) * W + z + 1] = v;
        zyx[((size_t)x * (YN + 1) + y + 1) * W + z + 1] = v;
      }
    }
  auto sat = [&](std::vector<T>& a, int H, int D) {
#pragma omp parallel for schedule(static)
    for (int d = 0; d < D; ++d) {
      T* pl = a.data() + (size_t)d * H * W;
      for (int h = 0; h < H; ++h)
        for (int w = 1; w < W; ++w) pl[(size_t)h * W + w] += pl[(size_t)h * W + w - 1];
      for (int h = 1; h < H; ++h)
        for (int w = 0; w < W; ++w) pl[(size_t)h * W + w] += pl[(size_t)(h - 1) * W + w];
    }
  };
  sat(zxy, XN + 1, YN);
  sat(zyx, YN + 1, XN);
}

// Per-ray set-up shared by the SAT and direct forward projectors
// (DD3_GPU.cu:874-923 / 957-961).
template <typename T>
struct Ray {
  bool xmarch;
  T fL, fR, fU, fD, Lx, Ly, Rx, Ry, Cx, Cy, zlo, zhi, constVal;
};

template <typename T>
Ray<T> ray_setup(const Geo<T>& g, int p, int u, int v) {
  const T c = g.cs[p], s = g.sn[p];
  const T sx = g.x0 * c - g.y0 * s, sy = g.x0 * s + g.y0 * c, sz = g.z0 + g.zsh[p];
  const T Cx = g.xds[u] * c - g.yds[u] * s, Cy = g.xds[u] * s + g.yds[u] * c;
  const T Lx = g.bxds[u] * c - g.byds[u] * s, Ly = g.bxds[u] * s + g.byds[u] * c;
  const T Rx = g.bxds[u + 1] * c - g.byds[u + 1] * s, Ry = g.bxds[u + 1] * s + g.byds[u + 1] * c;
  const T zlo = g.bzds[v] + g.zsh[p], zhi = g.bzds[v + 1] + g.zsh[p];
  T ddx = Cx - sx, ddy = Cy - sy, ddz = g.zds[v] + g.zsh[p] - sz;
  T nrm = std::sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
  ddx /= nrm; ddy /= nrm;
  Ray<T> r;
  r.Lx = Lx; r.Ly = Ly; r.Rx = Rx; r.Ry = Ry; r.Cx = Cx; r.Cy = Cy; r.zlo = zlo; r.zhi = zhi;
  if (std::fabs(c) <= std::fabs(s)) {
    r.xmarch = true;
    r.fL = (Ly - sy) / (Lx - sx);
    r.fR = (Ry - sy) / (Rx - sx);
    r.fU = (zhi - sz) / (Cx - sx);
    r.fD = (zlo - sz) / (Cx - sx);
    r.constVal = g.dx * g.dx * g.dz / std::fabs(ddx);
  } else {
    r.xmarch = false;
    r.fL = (Lx - sx) / (Ly - sy);
    r.fR = (Rx - sx) / (Ry - sy);
    r.fU = (zhi - sz) / (Cy - sy);
    r.fD = (zlo - sz) / (Cy - sy);
    r.constVal = g.dx * g.dx * g.dz / std::fabs(ddy);
  }
  return r;
}

// Footprint of ray r on slice `ii` in unpadded cell-centre index units:
// in-plane [tL, tR], z [tD, tU]; lenL/lenH in mm (DD3_GPU.cu:927-941 / 965-977).
template <typename T>
inline void footprint(const Geo<T>& g, const Ray<T>& r, int ii, T& tL, T& tR, T& tD, T& tU,
                      T& lenL, T& lenH) {
  T obj, realL, realR, realU, realD;
  if (r.xmarch) {
    obj = (T(ii) - g.ocx) * g.dx;
    realL = (obj - r.Lx) * r.fL + r.Ly;
    realR = (obj - r.Rx) * r.fR + r.Ry;
    realU = (obj - r.Cx) * r.fU + r.zhi;
    realD = (obj - r.Cx) * r.fD + r.zlo;
    tL = realL / g.dx + g.ocy;
    tR = realR / g.dx + g.ocy;
  } else {
    obj = (T(ii) - g.ocy) * g.dx;
    realL = (obj - r.Ly) * r.fL + r.Lx;
    realR = (obj - r.Ry) * r.fR + r.Rx;
    realU = (obj - r.Cy) * r.fU + r.zhi;
    realD = (obj - r.Cy) * r.fD + r.zlo;
    tL = realL / g.dx + g.ocx;
    tR = realR / g.dx + g.ocx;
  }
  tU = realU / g.dz + g.ocz;
  tD = realD / g.dz + g.ocz;
  lenL = realR - realL;
  lenH = realU - realD;
}

// ---------------- forward, SAT transcription (DD3_GPU.cu:844-990) ----------------
template <typename T, typename TO>
void fp_sat(const Geo<T>& g, const float* vol, TO* prj, int nthreads) {
  std::vector<T> zxy, zyx;
  volume_sats(g, vol, zxy, zyx);
  const int W = g.ZN + 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int p = 0; p < g.PN; ++p)
    for (int u = 0; u < g.DNU; ++u)
      for (int v = 0; v < g.DNV; ++v) {
        Ray<T> r = ray_setup(g, p, u, v);
        T summ = 0;
        const int NS = r.xmarch ? g.XN : g.YN;
        const int H = r.xmarch ? g.YN + 1 : g.XN + 1;
        const T* base = r.xmarch ? zyx.data() : zxy.data();
        for (int ii = 0; ii < NS; ++ii) {
          T tL, tR, tD, tU, lenL, lenH;
          footprint(g, r, ii, tL, tR, tD, tU, lenL, lenH);
          const T* pl = base + (size_t)ii * H * W;
          // texture coordinates carry +1 for the zero border
          T a = tex2(pl, W, H, tD + 1, tL + 1) + tex2(pl, W, H, tU + 1, tR + 1) -
                (tex2(pl, W, H, tD + 1, tR + 1) + tex2(pl, W, H, tU + 1, tL + 1));
          summ += a / (lenL * lenH);
        }
        prj[((size_t)p * g.DNU + u) * g.DNV + v] = TO(summ * r.constVal);
      }
}

// Exact overlap of [lo, hi] (sorted) with cell [k-0.5, k+0.5] clipped to [-0.5, N-0.5].
template <typename T>
inline void cell_range(T lo, T hi, int N, T& clo, T& chi, int& k0, int& k1) {
  clo = std::max(lo, T(-0.5));
  chi = std::min(hi, T(N) - T(0.5));
  if (!(clo < chi)) {
    k0 = 0;
    k1 = -1;
    return;
  }
  k0 = std::max(0, (int)std::floor(clo + T(0.5)));
  k1 = std::min(N - 1, (int)std::floor(chi + T(0.5)));
}

// ---------------- forward, direct restatement (no SAT) ----------------
template <typename T, typename TI, typename TO>
void fp_direct(const Geo<T>& g, const TI* vol, TO* prj, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int p = 0; p < g.PN; ++p)
    for (int u = 0; u < g.DNU; ++u)
      for (int v = 0; v < g.DNV; ++v) {
        Ray<T> r = ray_setup(g, p, u, v);
        T summ = 0;
        const int NS = r.xmarch ? g.XN : g.YN;
        const int NI = r.xmarch ? g.YN : g.XN;
        for (int ii = 0; ii < NS; ++ii) {
          T tL, tR, tD, tU, lenL, lenH;
          footprint(g, r, ii, tL, tR, tD, tU, lenL, lenH);
          T ylo = std::min(tL, tR), yhi = std::max(tL, tR);
          T zlo = std::min(tD, tU), zhi = std::max(tD, tU);
          T cyl, cyh, czl, czh;
          int j0, j1, k0, k1;
          cell_range(ylo, yhi, NI, cyl, cyh, j0, j1);
          cell_range(zlo, zhi, g.ZN, czl, czh, k0, k1);
          T acc = 0;
          for (int j = j0; j <= j1; ++j) {
            T wy = std::min(cyh, T(j) + T(0.5)) - std::max(cyl, T(j) - T(0.5));
            int x = r.xmarch ? ii : j, y = r.xmarch ? j : ii;
            T m = T(g.mask[y * g.XN + x]);
            if (m == T(0)) continue;
            const TI* col = vol + ((size_t)y * g.XN + x) * g.ZN;
            T s = 0;
            for (int k = k0; k <= k1; ++k) {
              T wz = std::min(czh, T(k) + T(0.5)) - std::max(czl, T(k) - T(0.5));
              s += wz * T(col[k]);
            }
            acc += wy * m * s;
          }
          summ += acc / std::fabs(lenL * lenH);
        }
        prj[((size_t)p * g.DNU + u) * g.DNV + v] = TO(summ * r.constVal);
      }
}

// ---------------- matched adjoint: exact transpose of fp_direct ----------------
template <typename T, typename TI>
void bp_adjoint(const Geo<T>& g, const TI* prj, double* vol) {
  const size_t NV = (size_t)g.XN * g.YN * g.ZN;
  std::fill(vol, vol + NV, 0.0);
  for (int p = 0; p < g.PN; ++p)
    for (int u = 0; u < g.DNU; ++u)
      for (int v = 0; v < g.DNV; ++v) {
        T yv = T(prj[((size_t)p * g.DNU + u) * g.DNV + v]);
        if (yv == T(0)) continue;
        Ray<T> r = ray_setup(g, p, u, v);
        const int NS = r.xmarch ? g.XN : g.YN;
        const int NI = r.xmarch ? g.YN : g.XN;
        for (int ii = 0; ii < NS; ++ii) {
          T tL, tR, tD, tU, lenL, lenH;
          footprint(g, r, ii, tL, tR, tD, tU, lenL, lenH);
          T ylo = std::min(tL, tR), yhi = std::max(tL, tR);
          T zlo = std::min(tD, tU), zhi = std::max(tD, tU);
          T cyl, cyh, czl, czh;
          int j0, j1, k0, k1;
          cell_range(ylo, yhi, NI, cyl, cyh, j0, j1);
          cell_range(zlo, zhi, g.ZN, czl, czh, k0, k1);
          T scale = yv * r.constVal / std::fabs(lenL * lenH);
          for (int j = j0; j <= j1; ++j) {
            T wy = std::min(cyh, T(j) + T(0.5)) - std::max(cyl, T(j) - T(0.5));
            int x = r.xmarch ? ii : j, y = r.xmarch ? j : ii;
            T m = T(g.mask[y * g.XN + x]);
            if (m == T(0)) continue;
            double* col = vol + ((size_t)y * g.XN + x) * g.ZN;
            for (int k = k0; k <= k1; ++k) {
              T wz = std::min(czh, T(k) + T(0.5)) - std::max(czl, T(k) - T(0.5));
              col[k] += double(scale * wy * m * wz);
            }
          }
        }
      }
}

// calDetParas (DD3_GPU.cu:1189-1217): (detCtrIdxU, detCtrIdxV, dbeta, ddv)
template <typename T>
void det_paras(const Geo<T>& g, T& cU, T& cV, T& dbeta, T& ddv) {
  const int DNU = g.DNU, DNV = g.DNV;
  ddv = (g.bzds[DNV] - g.bzds[0]) / T(DNV);
  cV = (-(g.bzds[0] - g.z0) / ddv) - T(0.5);
  auto nrm2 = [](T a, T b, T& oa, T& ob) {
    T n = std::sqrt(a * a + b * b);
    oa = a / n;
    ob = b / n;
  };
  T dirx, diry, lx, ly, rx, ry;
  nrm2(-g.x0, -g.y0, dirx, diry);
  nrm2(g.bxds[0] - g.x0, g.byds[0] - g.y0, lx, ly);
  nrm2(g.bxds[DNU] - g.x0, g.byds[DNU] - g.y0, rx, ry);
  dbeta = std::asin(lx * ry - ly * rx) / T(DNU);
  T minBeta = std::asin(dirx * ly - diry * lx);
  cU = -minBeta / dbeta - T(0.5);
}

// Per voxel-view shadow for the reference back projector (DD3_GPU.cu:1420-1464),
// in unpadded cell-centre detector index units.
template <typename T>
struct Shadow {
  T iL, iR, iU, iD, cosVal;
};

template <typename T>
inline Shadow<T> shadow(const Geo<T>& g, int p, T Px, T Py, T Pz, T S, T ib, T hz, T inv_sid,
                        T cU, T cV) {
  const T c = g.cs[p], s = g.sn[p];
  const T sx = g.x0 * c - g.y0 * s, sy = g.x0 * s + g.y0 * c, sz = g.z0 + g.zsh[p];
  const T dxv = Px - sx, dyv = Py - sy, dzv = Pz - sz;
  const T l2 = dxv * dxv + dyv * dyv;
  const T l = T(1) / std::sqrt(l2);
  Shadow<T> r;
  r.iU = (dzv + hz) * S * l + cV;
  r.iD = (dzv - hz) * S * l + cV;
  T alpha = std::asin((sy * dxv - sx * dyv) * inv_sid * l);
  T w = (std::fabs(sx) > std::fabs(sy)) ? dxv : dyv;
  T dA = w / l2 * g.dx * T(0.5);
  r.cosVal = g.dx / w * std::sqrt(l2 + dzv * dzv);
  r.iL = (alpha - dA) * ib + cU;
  r.iR = (alpha + dA) * ib + cU;
  return r;
}

template <typename T>
struct BackConsts {
  T S, ib, hz, inv_sid, cU, cV;
};

template <typename T>
BackConsts<T> back_consts(const Geo<T>& g) {
  BackConsts<T> k;
  T dbeta, ddv;
  det_paras(g, k.cU, k.cV, dbeta, ddv);
  T S2D = std::hypot(g.xds[0] - g.x0, g.yds[0] - g.y0);  // DD3_GPU.cu:1674
  k.S = S2D / ddv;
  k.ib = T(1) / dbeta;
  k.hz = g.dz * T(0.5);
  k.inv_sid = T(1) / std::sqrt(g.x0 * g.x0 + g.y0 * g.y0);
  return k;
}

// ---------------- back, SAT transcription (DD3_GPU.cu:1397-1475) ----------------
// genSAT_of_Projection (DD3_GPU.cu:1241-1323): per view, the sinogram with a zero row
// and column in front, integrated along v then u: sat[p][u'][v'] (W = DNV+1, H = DNU+1).
template <typename T>
std::vector<T> sino_sat(const Geo<T>& g, const float* prj) {
  const int DNU = g.DNU, DNV = g.DNV, W = DNV + 1, H = DNU + 1;
  std::vector<T> sat((size_t)g.PN * H * W, T(0));
  for (int p = 0; p < g.PN; ++p)
    for (int u = 0; u < DNU; ++u)
      for (int v = 0; v < DNV; ++v)
        sat[((size_t)p * H + u + 1) * W + v + 1] = T(prj[((size_t)p * DNU + u) * DNV + v]);
  for (int p = 0; p < g.PN; ++p) {
    T* pl = sat.data() + (size_t)p * H * W;
    for (int h = 0; h < H; ++h)
      for (int w = 1; w < W; ++w) pl[(size_t)h * W + w] += pl[(size_t)h * W + w - 1];
    for (int h = 1; h < H; ++h)
      for (int w = 0; w < W; ++w) pl[(size_t)h * W + w] += pl[(size_t)(h - 1) * W + w];
  }
  return sat;
}

template <typename T, typename TO>
void bp_sat(const Geo<T>& g, const float* prj, TO* vol, int nthreads) {
  const int DNU = g.DNU, DNV = g.DNV, W = DNV + 1, H = DNU + 1;
  const std::vector<T> sat = sino_sat(g, prj);
  const BackConsts<T> k = back_consts(g);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int y = 0; y < g.YN; ++y)
    for (int x = 0; x < g.XN; ++x) {
      TO* col = vol + ((size_t)y * g.XN + x) * g.ZN;
      if (g.mask[y * g.XN + x] != 1) {
        for (int z = 0; z < g.ZN; ++z) col[z] = TO(0);
        continue;
      }
      const T Px = (T(x) - g.ocx) * g.dx, Py = (T(y) - g.ocy) * g.dx;
      for (int z = 0; z < g.ZN; ++z) {
        const T Pz = (T(z) - g.ocz) * g.dz;
        T summ = 0;
        for (int p = 0; p < g.PN; ++p) {
          Shadow<T> sh = shadow(g, p, Px, Py, Pz, k.S, k.ib, k.hz, k.inv_sid, k.cU, k.cV);
          const T* pl = sat.data() + (size_t)p * H * W;
          summ += (-tex2(pl, W, H, sh.iD + 1, sh.iR + 1) - tex2(pl, W, H, sh.iU + 1, sh.iL + 1) +
                   tex2(pl, W, H, sh.iD + 1, sh.iL + 1) + tex2(pl, W, H, sh.iU + 1, sh.iR + 1)) *
                  sh.cosVal;
        }
        col[z] = TO(summ);
      }
    }
}

// ---------------- back, direct restatement ----------------
template <typename T, typename TI, typename TO>
void bp_direct(const Geo<T>& g, const TI* prj, TO* vol, int nthreads) {
  const BackConsts<T> k = back_consts(g);
  const int DNU = g.DNU, DNV = g.DNV;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int y = 0; y < g.YN; ++y)
    for (int x = 0; x < g.XN; ++x) {
      TO* col = vol + ((size_t)y * g.XN + x) * g.ZN;
      if (g.mask[y * g.XN + x] != 1) {
        for (int z = 0; z < g.ZN; ++z) col[z] = TO(0);
        continue;
      }
      const T Px = (T(x) - g.ocx) * g.dx, Py = (T(y) - g.ocy) * g.dx;
      for (int z = 0; z < g.ZN; ++z) {
        const T Pz = (T(z) - g.ocz) * g.dz;
        T summ = 0;
        for (int p = 0; p < g.PN; ++p) {
          Shadow<T> sh = shadow(g, p, Px, Py, Pz, k.S, k.ib, k.hz, k.inv_sid, k.cU, k.cV);
          T sgn = ((sh.iR - sh.iL) * (sh.iU - sh.iD) < 0) ? T(-1) : T(1);
          T ul, uh, vl, vh;
          int u0, u1, v0, v1;
          cell_range(std::min(sh.iL, sh.iR), std::max(sh.iL, sh.iR), DNU, ul, uh, u0, u1);
          cell_range(std::min(sh.iD, sh.iU), std::max(sh.iD, sh.iU), DNV, vl, vh, v0, v1);
          T acc = 0;
          for (int u = u0; u <= u1; ++u) {
            T wu = std::min(uh, T(u) + T(0.5)) - std::max(ul, T(u) - T(0.5));
            const TI* row = prj + ((size_t)p * DNU + u) * DNV;
            T s = 0;
            for (int v = v0; v <= v1; ++v) {
              T wv = std::min(vh, T(v) + T(0.5)) - std::max(vl, T(v) - T(0.5));
              s += wv * T(row[v]);
            }
            acc += wu * s;
          }
          summ += sgn * acc * sh.cosVal;
        }
        col[z] = TO(summ);
      }
    }
}

// ---------------- pseudo distance driven (prjMode 1 / 3 forward, 1 / 2 back) ----------------
// Linear-filtered fetch with border addressing (texels outside the array read 0) at
// texel-index coordinates (a, b) along W and H, i.e. the unnormalised texture coordinate
// minus 0.5.  f(w, h) returns texel (w, h).
template <typename T, typename F>
inline T lerp2_border(const F& f, int W, int H, T a, T b) {
  if (!(a > T(-1) && a < T(W) && b > T(-1) && b < T(H))) return T(0);
  const T fa = std::floor(a), fb = std::floor(b);
  const int i = (int)fa, j = (int)fb;
  const T al = a - fa, be = b - fb;
  auto at = [&](int w, int h) { return (w < 0 || w >= W || h < 0 || h >= H) ? T(0) : f(w, h); };
  return (T(1) - be) * ((T(1) - al) * at(i, j) + al * at(i + 1, j)) +
         be * ((T(1) - al) * at(i, j + 1) + al * at(i + 1, j + 1));
}

// DD3_GPU.cu:673-754 (kernel), :756-834 (host: vol *= mask; border-addressed linear
// texture of width ZN, height XN, depth YN).  One ray from the source to each detector
// cell centre, sampled where it crosses the centre plane of every x-slice (|cos| <=
// |sin|) or y-slice, times the path length per slice dx / |dir_x| (or / |dir_y|).
template <typename T>
void fp_pdd(const Geo<T>& g, const float* vol, T* prj) {
  const int XN = g.XN, ZN = g.ZN;
  auto voxel = [&](int x, int y, int z) {
    return T(vol[((size_t)y * XN + x) * ZN + z]) * T(g.mask[(size_t)y * XN + x]);
  };
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int p = 0; p < g.PN; ++p)
    for (int u = 0; u < g.DNU; ++u)
      for (int v = 0; v < g.DNV; ++v) {
        const T c = g.cs[p], s = g.sn[p];
        const T sx = g.x0 * c - g.y0 * s, sy = g.x0 * s + g.y0 * c, sz = g.z0 + g.zsh[p];
        const T Dx = g.xds[u] * c - g.yds[u] * s, Dy = g.xds[u] * s + g.yds[u] * c;
        const T Dz = g.zds[v] + g.zsh[p];
        T ex = Dx - sx, ey = Dy - sy, ez = Dz - sz;
        const T n = std::sqrt(ex * ex + ey * ey + ez * ez);
        ex /= n; ey /= n; ez /= n;
        T summ = 0, out;
        if (std::fabs(c) <= std::fabs(s)) {
          for (int ii = 0; ii < XN; ++ii) {
            const T obj = (T(ii) - g.ocx) * g.dx;
            const T yy = (obj - Dx) / ex * ey + Dy, zz = (obj - Dx) / ex * ez + Dz;
            summ += lerp2_border([&](int w, int h) { return voxel(ii, h, w); }, ZN, g.YN,
                                 zz / g.dz + g.ocz, yy / g.dx + g.ocy);
          }
          out = summ * g.dx / std::fabs(ex);
        } else {
          for (int jj = 0; jj < g.YN; ++jj) {
            const T obj = (T(jj) - g.ocy) * g.dx;
            const T xx = (obj - Dy) / ey * ex + Dx, zz = (obj - Dy) / ey * ez + Dz;
            summ += lerp2_border([&](int w, int h) { return voxel(h, jj, w); }, ZN, XN,
                                 zz / g.dz + g.ocz, xx / g.dx + g.ocx);
          }
          out = summ * g.dx / std::fabs(ey);
        }
        prj[((size_t)p * g.DNU + u) * g.DNV + v] = out;
      }
}

// DD3_GPU.cu:1478-1528 (kernel), host :1664-1700 (border-addressed linear texture of
// width DNV, height DNU, depth PN; constants of calDetParas).  Voxel driven: each view's
// sinogram is sampled where the ray from the source through the voxel centre meets the
// detector, weighted by dx |dir| / |dir_x| (or / |dir_y|, chosen by the larger component
// of the source position).  Voxels whose mask byte is not 1 are 0.
template <typename T>
void bp_pdd(const Geo<T>& g, const float* prj, T* vol) {
  const BackConsts<T> k = back_consts(g);
  const int DNU = g.DNU, DNV = g.DNV;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int y = 0; y < g.YN; ++y)
    for (int x = 0; x < g.XN; ++x) {
      T* col = vol + ((size_t)y * g.XN + x) * g.ZN;
      if (g.mask[y * g.XN + x] != 1) {
        for (int z = 0; z < g.ZN; ++z) col[z] = T(0);
        continue;
      }
      const T Px = (T(x) - g.ocx) * g.dx, Py = (T(y) - g.ocy) * g.dx;
      for (int z = 0; z < g.ZN; ++z) {
        const T Pz = (T(z) - g.ocz) * g.dz;
        T summ = 0;
        for (int p = 0; p < g.PN; ++p) {
          const T c = g.cs[p], s = g.sn[p];
          const T sx = g.x0 * c - g.y0 * s, sy = g.x0 * s + g.y0 * c, sz = g.z0 + g.zsh[p];
          const T dX = Px - sx, dY = Py - sy, dZ = Pz - sz;
          const T ddv = dX * dX + dY * dY;
          const T invl = T(1) / std::sqrt(ddv);
          const T iz = dZ * k.S * invl + k.cV;
          const T alpha = std::asin((sy * dX - sx * dY) * invl * k.inv_sid);
          T cosVal = (std::fabs(sx) >= std::fabs(sy)) ? std::fabs(T(1) / dX) : std::fabs(T(1) / dY);
          cosVal *= g.dx * std::sqrt(ddv + dZ * dZ);
          const T iu = alpha * k.ib + k.cU;
          const float* plane = prj + (size_t)p * DNU * DNV;
          summ += lerp2_border([&](int w, int h) { return T(plane[(size_t)h * DNV + w]); }, DNV,
                               DNU, iz, iu) *
                  cosVal;
        }
        col[z] = summ;
      }
    }
}

// ---------------- z-line back (prjMode 3; DD3_GPU.cu:1545-1647, host :1723-1737) --------
// Per voxel column, in blocks of 16 slices (only ZN / 16 whole blocks; the slices of a
// last partial block stay 0): for every view, 17 terms at the lower faces of the slices
// lp*16 + j, j = 0..16, each the sinogram integral over the fan-angle range [tL, tR] of
// the column's y- (or x-) extent and the rows below the face, times the signed cell
// factor dx |d_xy| / d_x (or / d_y) and hypot(d_z, |d_xy|) / |d_xy| evaluated at the slice
// centre; a slice gets the difference of the sums at its upper and lower faces.
template <typename T>
void bp_zline(const Geo<T>& g, const float* prj, T* vol) {
  const int DNU = g.DNU, DNV = g.DNV, W = DNV + 1, H = DNU + 1;
  const std::vector<T> sat = sino_sat(g, prj);
  const BackConsts<T> k = back_consts(g);
  const int LPs = g.ZN >> 4;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int y = 0; y < g.YN; ++y)
    for (int x = 0; x < g.XN; ++x) {
      T* col = vol + ((size_t)y * g.XN + x) * g.ZN;
      for (int z = 0; z < g.ZN; ++z) col[z] = T(0);
      if (g.mask[y * g.XN + x] != 1) continue;
      const T Px = (T(x) - g.ocx) * g.dx, Py = (T(y) - g.ocy) * g.dx;
      for (int lp = 0; lp < LPs; ++lp) {
        T summ[17] = {};
        const T minObj = (T(lp * 16) - g.ocz) * g.dz;
        for (int p = 0; p < g.PN; ++p) {
          const T c = g.cs[p], s = g.sn[p];
          const T sx = g.x0 * c - g.y0 * s, sy = g.x0 * s + g.y0 * c, sz = g.z0 + g.zsh[p];
          const T dX = Px - sx, dY = Py - sy;
          const T s2v = std::hypot(dX, dY);
          T lx, ly, rx, ry, cosVal;
          if (std::fabs(c) <= std::fabs(s)) {
            lx = dX; ly = dY - T(0.5) * g.dx;
            rx = dX; ry = dY + T(0.5) * g.dx;
            cosVal = g.dx * s2v / dX;
          } else {
            lx = dX + T(0.5) * g.dx; ly = dY;
            rx = dX - T(0.5) * g.dx; ry = dY;
            cosVal = g.dx * s2v / dY;
          }
          const T nl = std::sqrt(lx * lx + ly * ly), nr = std::sqrt(rx * rx + ry * ry);
          const T tL = std::asin((sy * lx - sx * ly) / nl * k.inv_sid) * k.ib + k.cU;
          const T tR = std::asin((sy * rx - sx * ry) / nr * k.inv_sid) * k.ib + k.cU;
          const T curang = k.S / s2v;
          const T* pl = sat.data() + (size_t)p * H * W;
          for (int j = 0; j <= 16; ++j) {
            const T dirZ = minObj + T(j) * g.dz - sz;
            const T w = std::hypot(dirZ, s2v) / s2v;
            const T tD = (dirZ - T(0.5) * g.dz) * curang + k.cV;
            summ[j] += (tex2(pl, W, H, tD + 1, tR + 1) - tex2(pl, W, H, tD + 1, tL + 1)) *
                       cosVal * w;
          }
        }
        for (int j = 0; j < 16; ++j) col[lp * 16 + j] = summ[j + 1] - summ[j];
      }
    }
}

// ---------------- FP64 branchless (prjMode 2 forward) ----------------
// Point-sampled integral image with clamp addressing and explicit bilinear weights
// between floor and ceil of the coordinates (DD3_GPU.cu:369-392, bilerp of
// utilities.cuh:245-254): first along H at both W rows, then along W.
template <typename T>
inline T sat_point(const T* pl, int W, int H, T a, T b) {
  a = std::min(std::max(a, T(-1)), T(W));
  b = std::min(std::max(b, T(-1)), T(H));
  const T fa = std::floor(a), ca = std::ceil(a), fb = std::floor(b), cb = std::ceil(b);
  auto at = [&](T w, T h) {
    const int wi = std::min(std::max((int)w, 0), W - 1), hi = std::min(std::max((int)h, 0), H - 1);
    return pl[(size_t)hi * W + wi];
  };
  const T t1 = b - fb, t2 = a - fa;
  const T vv0 = at(fa, fb) * (T(1) - t1) + at(fa, cb) * t1;
  const T vv1 = at(ca, fb) * (T(1) - t1) + at(ca, cb) * t1;
  return vv0 * (T(1) - t2) + vv1 * t2;
}

// DD3_GPU.cu:255-466 (kernel), :468-660 (host): the branchless footprint model with the
// texel coordinates real / d + objCntIdx + 1 sampled without the half-texel offset of
// linear filtering.
template <typename T>
void fp_dp(const Geo<T>& g, const float* vol, T* prj) {
  std::vector<T> zxy, zyx;
  volume_sats(g, vol, zxy, zyx);
  const int W = g.ZN + 1;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int p = 0; p < g.PN; ++p)
    for (int u = 0; u < g.DNU; ++u)
      for (int v = 0; v < g.DNV; ++v) {
        Ray<T> r = ray_setup(g, p, u, v);
        T summ = 0;
        const int NS = r.xmarch ? g.XN : g.YN;
        const int H = r.xmarch ? g.YN + 1 : g.XN + 1;
        const T* base = r.xmarch ? zyx.data() : zxy.data();
        for (int ii = 0; ii < NS; ++ii) {
          T tL, tR, tD, tU, lenL, lenH;
          footprint(g, r, ii, tL, tR, tD, tU, lenL, lenH);
          const T* pl = base + (size_t)ii * H * W;
          const T a = sat_point(pl, W, H, tD + 1, tL + 1) + sat_point(pl, W, H, tU + 1, tR + 1) -
                      sat_point(pl, W, H, tD + 1, tR + 1) - sat_point(pl, W, H, tU + 1, tL + 1);
          summ += a / (lenL * lenH);
        }
        prj[((size_t)p * g.DNU + u) * g.DNV + v] = summ * r.constVal;
      }
}

// Geometry restricted to the views s, s+osNum, s+2*osNum, ... (os_sart.m:15/23,
// zero-based subset index s).
template <typename T>
Geo<T> subset_geo(const Geo<T>& g, int s, int osNum) {
  Geo<T> h = g;
  h.cs.clear(); h.sn.clear(); h.zsh.clear();
  for (int p = s; p < g.PN; p += osNum) {
    h.cs.push_back(g.cs[p]);
    h.sn.push_back(g.sn[p]);
    h.zsh.push_back(g.zsh[p]);
  }
  h.PN = (int)h.cs.size();
  return h;
}

// Back projector selected by `kind`: 0 = reference arithmetic (DD3_GPU.cu:1397-1475,
// direct restatement), 1 = matched adjoint of the forward projector.
template <typename TI>
void back_any(const Geo<double>& g, const TI* prj, double* vol, int kind) {
  if (kind == 1)
    bp_adjoint(g, prj, vol);
  else
    bp_direct(g, prj, vol, 0);
}

// os_sart.m:13-44 in FP64 (relaxation 1, no clamp, guarded divisions).
// trace (may be NULL) receives the iterate after every sweep: trace[it * NV + i].
void os_sart(const Geo<double>& g, double* x, const float* b, int osNum, int iterNum, int kind,
             double* resid_out, double* trace) {
  const size_t NV = (size_t)g.XN * g.YN * g.ZN, NR = (size_t)g.DNU * g.DNV;
  std::vector<Geo<double>> sg;
  std::vector<std::vector<double>> rowSum(osNum), colSum(osNum);
  std::vector<double> ones(NV, 1.0);
  for (int s = 0; s < osNum; ++s) {
    sg.push_back(subset_geo(g, s, osNum));
    rowSum[s].resize(NR * sg[s].PN);
    fp_direct(sg[s], ones.data(), rowSum[s].data(), 0);
    std::vector<double> onesS(NR * sg[s].PN, 1.0);
    colSum[s].resize(NV);
    back_any(sg[s], onesS.data(), colSum[s].data(), kind);
  }
  std::vector<double> tp, tm(NV);
  for (int it = 0; it < iterNum; ++it) {
    double rr = 0, bb = 0;
    for (int s = 0; s < osNum; ++s) {
      const int P = sg[s].PN;
      tp.assign(NR * P, 0.0);
      fp_direct(sg[s], x, tp.data(), 0);
      for (int k = 0; k < P; ++k) {
        const int p = s + k * osNum;
        for (size_t r = 0; r < NR; ++r) {
          double bv = b[(size_t)p * NR + r];
          double d = bv - tp[(size_t)k * NR + r];
          rr += d * d;
          bb += bv * bv;
          double rs = rowSum[s][(size_t)k * NR + r];
          tp[(size_t)k * NR + r] = (rs != 0.0) ? d / rs : d;
        }
      }
      back_any(sg[s], tp.data(), tm.data(), kind);
      for (size_t i = 0; i < NV; ++i) {
        double cs = colSum[s][i];
        x[i] += (cs != 0.0) ? tm[i] / cs : tm[i];
      }
    }
    if (resid_out) resid_out[it] = std::sqrt(rr / bb);  // residual seen during the sweep
    if (trace) std::copy(x, x + NV, trace + (size_t)it * NV);
  }
}

// conjgrad.m:4-27 in FP64 with A = Back(Proj(.)), b = Back(prj).
void conjgrad(const Geo<double>& g, double* x, const float* prj, int iterNum, int kind,
              double* resid_out, double* trace) {
  const size_t NV = (size_t)g.XN * g.YN * g.ZN, NS = (size_t)g.DNU * g.DNV * g.PN;
  std::vector<double> sino(NS), b(NV), r0(NV), p0(NV), t(NV);
  auto A = [&](const double* in, double* out) {
    fp_direct(g, in, sino.data(), 0);
    back_any(g, sino.data(), out, kind);
  };
  back_any(g, prj, b.data(), kind);
  A(x, t.data());
  for (size_t i = 0; i < NV; ++i) r0[i] = b[i] - t[i];
  p0 = r0;
  for (int it = 0; it < iterNum; ++it) {
    double tmp1 = 0, pap = 0;
    for (size_t i = 0; i < NV; ++i) tmp1 += r0[i] * r0[i];
    A(p0.data(), t.data());
    for (size_t i = 0; i < NV; ++i) pap += p0[i] * t[i];
    const double alpha = tmp1 / pap;
    double r1r1 = 0;
    for (size_t i = 0; i < NV; ++i) {
      x[i] += alpha * p0[i];
      r0[i] -= alpha * t[i];
      r1r1 += r0[i] * r0[i];
    }
    const double beta = r1r1 / tmp1;
    for (size_t i = 0; i < NV; ++i) p0[i] = r0[i] + beta * p0[i];
    if (resid_out) resid_out[it] = std::sqrt(r1r1);
    if (trace) std::copy(x, x + NV, trace + (size_t)it * NV);
  }
}

}  // namespace

// ================================ C ABI ================================
// Argument lists follow DD3Proj_gpu / DD3Back_gpu (iterativeRecon/DD3_GPU.cu:1132-1142,
// 1766-1775) with the output typed by the routine.

#define ORC_GEO_ARGS                                                                         \
  float x0, float y0, float z0, int DNU, int DNV, const float *xds, const float *yds,         \
      const float *zds, float cx, float cy, float cz, const float *angs, const float *zpos, \
      int PN, int XN, int YN, int ZN
#define ORC_GEO_PASS x0, y0, z0, DNU, DNV, xds, yds, zds, cx, cy, cz, angs, zpos, PN, XN, YN, ZN

extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_boundaries_f64(int nb, const double* c, double* b) { boundaries<double>(nb, c, b); }

void orc_det_paras_f64(ORC_GEO_ARGS, float dx, float dz, double* out4) {
  Geo<double> g = make_geo<double>(ORC_GEO_PASS, dx, dz, nullptr);
  det_paras(g, out4[0], out4[1], out4[2], out4[3]);
}

void orc_fp_sat_f64(ORC_GEO_ARGS, const float* vol, double* prj, float dx, float dz,
                    const uint8_t* mask) {
  fp_sat(make_geo<double>(ORC_GEO_PASS, dx, dz, mask), vol, prj, 0);
}

void orc_fp_direct_f64(ORC_GEO_ARGS, const float* vol, double* prj, float dx, float dz,
                       const uint8_t* mask) {
  fp_direct(make_geo<double>(ORC_GEO_PASS, dx, dz, mask), vol, prj, 0);
}

void orc_bp_sat_f64(ORC_GEO_ARGS, double* vol, const float* prj, float dx, float dz,
                    const uint8_t* mask) {
  bp_sat(make_geo<double>(ORC_GEO_PASS, dx, dz, mask), prj, vol, 0);
}

void orc_bp_direct_f64(ORC_GEO_ARGS, double* vol, const float* prj, float dx, float dz,
                       const uint8_t* mask) {
  bp_direct(make_geo<double>(ORC_GEO_PASS, dx, dz, mask), prj, vol, 0);
}

void orc_bp_adjoint_f64(ORC_GEO_ARGS, double* vol, const float* prj, float dx, float dz,
                        const uint8_t* mask) {
  bp_adjoint(make_geo<double>(ORC_GEO_PASS, dx, dz, mask), prj, vol);
}

void orc_fp_pdd_f64(ORC_GEO_ARGS, const float* vol, double* prj, float dx, float dz,
                    const uint8_t* mask) {
  fp_pdd(make_geo<double>(ORC_GEO_PASS, dx, dz, mask), vol, prj);
}

void orc_bp_pdd_f64(ORC_GEO_ARGS, double* vol, const float* prj, float dx, float dz,
                    const uint8_t* mask) {
  bp_pdd(make_geo<double>(ORC_GEO_PASS, dx, dz, mask), prj, vol);
}

void orc_bp_zline_f64(ORC_GEO_ARGS, double* vol, const float* prj, float dx, float dz,
                      const uint8_t* mask) {
  bp_zline(make_geo<double>(ORC_GEO_PASS, dx, dz, mask), prj, vol);
}

void orc_fp_dp_f64(ORC_GEO_ARGS, const float* vol, double* prj, float dx, float dz,
                   const uint8_t* mask) {
  fp_dp(make_geo<double>(ORC_GEO_PASS, dx, dz, mask), vol, prj);
}

// FP32 instantiation of the SAT transcription: the CPU baseline.
void orc_fp_sat_f32(ORC_GEO_ARGS, const float* vol, float* prj, float dx, float dz,
                    const uint8_t* mask, int nthreads) {
  fp_sat(make_geo<float>(ORC_GEO_PASS, dx, dz, mask), vol, prj, nthreads);
}

void orc_bp_sat_f32(ORC_GEO_ARGS, float* vol, const float* prj, float dx, float dz,
                    const uint8_t* mask, int nthreads) {
  bp_sat(make_geo<float>(ORC_GEO_PASS, dx, dz, mask), prj, vol, nthreads);
}

}  // extern "C"

extern "C" {

// x (in/out, FP64 volume) is updated in place; resid[iterNum] receives the
// relative residual seen during each sweep; trace (NULL, or iterNum volumes) every
// iterate.  kind: 0 reference bp, 1 matched.
void orc_os_sart_f64(ORC_GEO_ARGS, double* x, const float* prj, float dx, float dz,
                     const uint8_t* mask, int osNum, int iterNum, int kind, double* resid,
                     double* trace) {
  os_sart(make_geo<double>(ORC_GEO_PASS, dx, dz, mask), x, prj, osNum, iterNum, kind, resid,
          trace);
}

void orc_conjgrad_f64(ORC_GEO_ARGS, double* x, const float* prj, float dx, float dz,
                      const uint8_t* mask, int iterNum, int kind, double* resid,
                      double* trace) {
  conjgrad(make_geo<double>(ORC_GEO_PASS, dx, dz, mask), x, prj, iterNum, kind, resid, trace);
}

// FP64 forward projection of an FP64 volume (direct form), for residuals of
// FP64 iterates.
void orc_fp_direct_f64_in64(ORC_GEO_ARGS, const double* vol, double* prj, float dx, float dz,
                            const uint8_t* mask) {
  fp_direct(make_geo<double>(ORC_GEO_PASS, dx, dz, mask), vol, prj, 0);
}

}  // extern "C"
