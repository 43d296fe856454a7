// Host side of the library: geometry tables (FP64 arithmetic, FP32 storage),
// the C ABI, and the OS-SART / CG drivers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "ctr_internal.hpp"

namespace ctr {

namespace {
thread_local std::string g_last_error;
constexpr double kTwoPi = 6.283185307179586;  // utilities.cuh:75
// allocation-failure injection: the allocation that finds 0 here fails (and disarms it);
// -1: off
std::atomic<long long> g_alloc_countdown{-1};
}  // namespace

const char* last_error_cstr() { return g_last_error.c_str(); }

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

hipError_t dev_malloc(void** p, size_t bytes) {
  long long n = g_alloc_countdown.load();
  while (n >= 0) {
    if (g_alloc_countdown.compare_exchange_weak(n, n == 0 ? -1 : n - 1)) {
      if (n == 0) {
        *p = nullptr;
        return hipErrorOutOfMemory;
      }
      break;
    }
  }
  return hipMalloc(p, std::max<size_t>(bytes, 1));
}

int ensure_bytes(void** p, size_t* have, size_t need, const char* what) {
  if (*have >= need && *p) return CTR_OK;
  (void)hipFree(*p);
  *p = nullptr;
  *have = 0;
  if (dev_malloc(p, need) != hipSuccess) {
    *p = nullptr;
    return fail(CTR_ERR_NOMEM, std::string("cannot allocate ") + what);
  }
  *have = need;
  return CTR_OK;
}

namespace {

// DD3Boundaries (iterativeRecon/utilities.cuh:340-361) in FP64.
std::vector<double> boundaries(const float* c, int n) {
  std::vector<double> b(n + 1);
  if (n + 1 >= 3) {
    b[0] = 1.5 * c[0] - 0.5 * c[1];
    for (int i = 1; i < n; ++i) b[i] = 0.5 * c[i - 1] + 0.5 * c[i];
    b[n] = 1.5 * c[n - 1] - 0.5 * c[n - 2];
  } else {
    b[0] = c[0] - 0.5;
    b[1] = c[0] + 0.5;
  }
  return b;
}

double regularize(double a) {  // utilities.cuh:257-265
  while (a >= kTwoPi) a -= kTwoPi;
  while (a < 0) a += kTwoPi;
  return a;
}

bool finite_array(const float* a, int n) {
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(a[i])) return false;
  return true;
}

template <typename T>
int upload(T** dst, const std::vector<T>& src) {
  if (dev_malloc((void**)dst, sizeof(T) * std::max<size_t>(src.size(), 1)) != hipSuccess) {
    *dst = nullptr;
    return fail(CTR_ERR_NOMEM, "cannot allocate the plan's geometry tables");
  }
  CTR_HIP(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
  return CTR_OK;
}

int validate_geometry(const ctr_geometry* g) {
  if (!g) return fail(CTR_ERR_ARG, "geometry is NULL");
  if (g->DNU < 1 || g->DNV < 1 || g->PN < 1 || g->XN < 1 || g->YN < 1 || g->ZN < 1)
    return fail(CTR_ERR_ARG, "all sizes (DNU, DNV, PN, XN, YN, ZN) must be >= 1");
  if (!g->xds || !g->yds || !g->zds || !g->angles || !g->zshifts)
    return fail(CTR_ERR_ARG, "detector or view arrays are NULL");
  if (!(g->dx > 0.f) || !(g->dz > 0.f)) return fail(CTR_ERR_ARG, "dx and dz must be > 0");
  if ((long long)g->XN * g->YN * g->ZN > (1LL << 31) - 1)
    return fail(CTR_ERR_ARG, "volume larger than 2^31-1 voxels");
  if ((long long)g->DNU * g->DNV > (1LL << 30))
    return fail(CTR_ERR_ARG, "detector frame too large");
  if (!finite_array(g->xds, g->DNU) || !finite_array(g->yds, g->DNU) ||
      !finite_array(g->zds, g->DNV) || !finite_array(g->angles, g->PN) ||
      !finite_array(g->zshifts, g->PN))
    return fail(CTR_ERR_NUMERIC, "non-finite geometry value");
  for (int v = 1; v < g->DNV; ++v)
    if (!(g->zds[v] > g->zds[v - 1]))
      return fail(CTR_ERR_UNSUPPORTED, "detector row positions zds must be strictly increasing");
  return CTR_OK;
}

}  // namespace

// Builds every table of the plan in FP64 on the host.
// half: 0, or 0.5 to move the footprints of the forward projector half a cell along the
// in-plane axis across the rays and along z (the sampling of the reference's FP64
// branchless projector, DD3_GPU.cu:350-360 / 414-422).
static int build_tables(const ctr_geometry* g, Plan& P, double half = 0.0) {
  const int DNU = g->DNU, DNV = g->DNV, PN = g->PN;
  const double x0 = g->x0, y0 = g->y0, z0 = g->z0, dx = g->dx, dz = g->dz;
  Dims& d = P.d;
  d.XN = g->XN; d.YN = g->YN; d.ZN = g->ZN; d.DNU = DNU; d.DNV = DNV; d.PN = PN;
  // objCntIdx, DD3_GPU.cu:1029-1031
  const double ocx = (g->XN - 1.0) * 0.5 - g->imgXCenter / dx;
  const double ocy = (g->YN - 1.0) * 0.5 - g->imgYCenter / dx;
  const double ocz = (g->ZN - 1.0) * 0.5 - g->imgZCenter / dz;
  d.ocx = (float)ocx; d.ocy = (float)ocy; d.ocz = (float)ocz;
  d.dx = (float)dx; d.dz = (float)dz;
  const std::vector<double> bx = boundaries(g->xds, DNU), by = boundaries(g->yds, DNU),
                            bz = boundaries(g->zds, DNV);

  // rows
  std::vector<float> Bz(DNV + 1), h2(DNV), invdB(DNV);
  for (int v = 0; v <= DNV; ++v) Bz[v] = (float)((bz[v] - z0) / dz);
  for (int v = 0; v < DNV; ++v) {
    const double h = g->zds[v] - z0;
    h2[v] = (float)(h * h);
    invdB[v] = (float)(dz / (bz[v + 1] - bz[v]));
  }
  d.B0 = Bz[0];
  d.dB = (float)((bz[DNV] - bz[0]) / dz / DNV);
  d.dBmin = Bz[1] - Bz[0];
  for (int v = 1; v < DNV; ++v) d.dBmin = std::min(d.dBmin, Bz[v + 1] - Bz[v]);

  // fan angles of the column boundaries seen from the source (rotation invariant)
  const double sid = std::sqrt(x0 * x0 + y0 * y0);
  auto fan = [&](double px, double py) {
    const double ex = px - x0, ey = py - y0, l = std::sqrt(ex * ex + ey * ey);
    return std::asin((y0 * ex - x0 * ey) / (sid * l));
  };
  d.beta0 = (float)fan(bx[0], by[0]);
  d.dbeta_g = (float)((fan(bx[DNU], by[DNU]) - fan(bx[0], by[0])) / DNU);
  if (!(std::fabs(d.dbeta_g) > 0.f))
    return fail(CTR_ERR_UNSUPPORTED, "detector columns subtend no fan angle");

  // reference back-projector constants (DD3_GPU.cu:1428-1437, calDetParas :1189-1217)
  {
    const double ddv = (bz[DNV] - bz[0]) / DNV;
    const double cV = (-(bz[0] - z0) / ddv) - 0.5;
    auto nrm = [](double a, double b, double& oa, double& ob) {
      const double n = std::sqrt(a * a + b * b);
      oa = a / n;
      ob = b / n;
    };
    double dirx, diry, lx, ly, rx, ry;
    nrm(-x0, -y0, dirx, diry);
    nrm(bx[0] - x0, by[0] - y0, lx, ly);
    nrm(bx[DNU] - x0, by[DNU] - y0, rx, ry);
    const double dbeta = std::asin(lx * ry - ly * rx) / DNU;
    const double minBeta = std::asin(dirx * ly - diry * lx);
    const double cU = -minBeta / dbeta - 0.5;
    const double S2D = std::hypot((double)g->xds[0] - x0, (double)g->yds[0] - y0);
    P.bc.S = (float)(S2D / ddv);
    P.bc.ib = (float)(1.0 / dbeta);
    P.bc.hz = (float)(dz * 0.5);
    P.bc.inv_sid = (float)(1.0 / sid);
    P.bc.cU = (float)cU;
    P.bc.cV = (float)cV;
    P.bc.dx = (float)dx;
    P.bc.dz = (float)dz;
    P.bc64.ib = 1.0 / dbeta;
    P.bc64.cU = cU;
    P.bc64.inv_sid = 1.0 / sid;
    P.bc64.dx = dx;
    P.bc64.ocx = ocx;
    P.bc64.ocy = ocy;
    P.bc64.S = S2D / ddv;
    P.bc64.cV = cV;
    P.bc64.ocz = ocz;
  }

  // per-view tables
  std::vector<ViewP> views(PN);
  std::vector<double2> src64(PN), rot64(PN), det64(DNU);
  std::vector<double> sz64(PN);
  std::vector<float> zc(DNV);
  for (int u = 0; u < DNU; ++u) det64[u] = make_double2(g->xds[u], g->yds[u]);
  for (int v = 0; v < DNV; ++v) zc[v] = (float)((g->zds[v] - z0) / dz);
  std::vector<float2> bnd((size_t)PN * (DNU + 1));
  std::vector<double2> bnd64((size_t)PN * (DNU + 1));
  std::vector<float4> cols((size_t)PN * DNU);
  for (int p = 0; p < PN; ++p) {
    const double th = regularize((double)g->angles[p]);
    const double c = std::cos(th), s = std::sin(th);
    const double sx = x0 * c - y0 * s, sy = x0 * s + y0 * c, sz = z0 + g->zshifts[p];
    ViewP& vp = views[p];
    vp.sx = (float)sx; vp.sy = (float)sy; vp.sz = (float)sz;
    src64[p] = make_double2(sx, sy);
    rot64[p] = make_double2(c, s);
    sz64[p] = sz;
    const double az = sz / dz + ocz + half;
    vp.az = (float)az;
    vp.azi = (int)std::floor(az + 0.5);
    vp.azf = (float)(az - vp.azi);
    vp.pad0 = vp.pad1 = vp.pad2 = 0;
    vp.xmarch = std::fabs(c) <= std::fabs(s) ? 1 : 0;  // DD3_GPU.cu:914
    vp.wx = std::fabs(sx) > std::fabs(sy) ? 1 : 0;  // DD3_GPU.cu:1453
    const int NS = vp.xmarch ? g->XN : g->YN;
    const double mid = 0.5 * (NS - 1);
    for (int b = 0; b <= DNU; ++b) {
      const double Bx = bx[b] * c - by[b] * s, By = bx[b] * s + by[b] * c;
      double f, a;
      if (vp.xmarch) {
        f = (By - sy) / (Bx - sx);
        a = ((-ocx * dx - Bx) * f + By) / dx + ocy + half;
      } else {
        f = (Bx - sx) / (By - sy);
        a = ((-ocy * dx - By) * f + Bx) / dx + ocx + half;
      }
      bnd[(size_t)p * (DNU + 1) + b] = make_float2((float)a, (float)f);
      bnd64[(size_t)p * (DNU + 1) + b] = make_double2(a, f);
    }
    {
      const float2 b0 = bnd[(size_t)p * (DNU + 1)], bN = bnd[(size_t)p * (DNU + 1) + DNU];
      vp.bdir = ((bN.x + mid * bN.y) >= (b0.x + mid * b0.y)) ? 1 : -1;
    }
    for (int u = 0; u < DNU; ++u) {
      const double Dx = (double)g->xds[u] * c - (double)g->yds[u] * s;
      const double Dy = (double)g->xds[u] * s + (double)g->yds[u] * c;
      const double ex = Dx - sx, ey = Dy - sy;
      double r0, r1, Q;
      if (vp.xmarch) {
        r0 = (-ocx * dx - sx) / ex;
        r1 = dx / ex;
        Q = dx / std::fabs(ex);
      } else {
        r0 = (-ocy * dx - sy) / ey;
        r1 = dx / ey;
        Q = dx / std::fabs(ey);
      }
      cols[(size_t)p * DNU + u] = make_float4((float)r0, (float)r1, (float)(ex * ex + ey * ey),
                                              (float)Q);
    }
  }
  P.h_views = views;
  {
    HostGeo& h = P.hg;
    h.sz.resize(PN);
    for (int p = 0; p < PN; ++p) h.sz[p] = z0 + g->zshifts[p];
    h.mask.assign((size_t)g->XN * g->YN, 1);
    if (g->mask)
      for (size_t i = 0; i < h.mask.size(); ++i) h.mask[i] = g->mask[i] != 0;
    h.ocx = ocx; h.ocy = ocy; h.ocz = ocz; h.dx = dx; h.dz = dz;
    h.S2D = std::hypot((double)g->xds[0] - x0, (double)g->yds[0] - y0);
    h.bz0 = bz[0] - z0;
    h.bzN = bz[DNV] - z0;
  }

  std::vector<uint8_t> mask((size_t)g->XN * g->YN, 1);
  if (g->mask) std::memcpy(mask.data(), g->mask, mask.size());

  // per-slice extent of the non-zero mask: x-slices (cells along y), then y-slices
  std::vector<float2> mext(g->XN + g->YN);
  for (int x = 0; x < g->XN; ++x) {
    int lo = g->YN, hi = -1;
    for (int y = 0; y < g->YN; ++y)
      if (mask[(size_t)y * g->XN + x]) { lo = std::min(lo, y); hi = std::max(hi, y); }
    mext[x] = (hi >= lo) ? make_float2(lo - 0.5f, hi + 0.5f) : make_float2(0.f, -1.f);
  }
  for (int y = 0; y < g->YN; ++y) {
    int lo = g->XN, hi = -1;
    for (int x = 0; x < g->XN; ++x)
      if (mask[(size_t)y * g->XN + x]) { lo = std::min(lo, x); hi = std::max(hi, x); }
    mext[g->XN + y] = (hi >= lo) ? make_float2(lo - 0.5f, hi + 0.5f) : make_float2(0.f, -1.f);
  }

  std::vector<float> maskf(mask.begin(), mask.end());

  int rc;
  if ((rc = upload(&P.mext, mext))) return rc;
  if ((rc = upload(&P.maskf, maskf))) return rc;
  if ((rc = upload(&P.views, views))) return rc;
  if ((rc = upload(&P.src64, src64))) return rc;
  if ((rc = upload(&P.rot64, rot64))) return rc;
  if ((rc = upload(&P.sz64, sz64))) return rc;
  if ((rc = upload(&P.det64, det64))) return rc;
  if ((rc = upload(&P.zc, zc))) return rc;
  if ((rc = upload(&P.bnd, bnd))) return rc;
  if ((rc = upload(&P.bnd64, bnd64))) return rc;
  if ((rc = upload(&P.cols, cols))) return rc;
  if ((rc = upload(&P.Bz, Bz))) return rc;
  if ((rc = upload(&P.h2, h2))) return rc;
  if ((rc = upload(&P.invdB, invdB))) return rc;
  if ((rc = upload(&P.mask, mask))) return rc;
  return CTR_OK;
}

static void free_plan(Plan* P) {
  if (!P) return;
  (void)hipFree(P->views);
  (void)hipFree(P->src64);
  (void)hipFree(P->rot64);
  (void)hipFree(P->sz64);
  (void)hipFree(P->zl_cpre);
  (void)hipFree(P->det64);
  (void)hipFree(P->zc);
  (void)hipFree(P->bnd);
  (void)hipFree(P->bnd64);
  (void)hipFree(P->cols);
  (void)hipFree(P->Bz);
  (void)hipFree(P->h2);
  (void)hipFree(P->invdB);
  (void)hipFree(P->mask);
  (void)hipFree(P->mext);
  (void)hipFree(P->maskf);
  (void)hipFree(P->scratch_prj);
  (void)hipFree(P->sat_vol);
  (void)hipFree(P->sat_sino);
  (void)hipFree(P->partials);
  for (void* b : P->cg) (void)hipFree(b);
  (void)hipFree(P->sums2);
  if (P->stream) (void)hipStreamDestroy(P->stream);
  delete P;
}

static int ensure_scratch_prj(Plan& P, size_t views) {
  if (P.scratch_prj_views >= views) return CTR_OK;
  (void)hipFree(P.scratch_prj);
  P.scratch_prj = nullptr;
  P.scratch_prj_views = 0;
  const size_t bytes = views * (size_t)P.d.DNU * P.d.DNV * sizeof(float);
  if (dev_malloc((void**)&P.scratch_prj, bytes) != hipSuccess)
    return fail(CTR_ERR_NOMEM, "cannot allocate sinogram scratch");
  P.scratch_prj_views = views;
  return CTR_OK;
}

static int ensure_partials(Plan& P, size_t n) {
  if (P.partials_n >= n) return CTR_OK;
  (void)hipFree(P.partials);
  P.partials = nullptr;
  P.partials_n = 0;
  if (dev_malloc((void**)&P.partials, n * sizeof(double)) != hipSuccess)
    return fail(CTR_ERR_NOMEM, "cannot allocate reduction scratch");
  P.partials_n = n;
  return CTR_OK;
}

static int check_views(const Plan& P, int vb, int vs, int vc) {
  if (vc < 0 || vb < 0 || vs < 1) return fail(CTR_ERR_ARG, "bad view range");
  if (vc > 0 && (long long)vb + (long long)(vc - 1) * vs >= P.d.PN)
    return fail(CTR_ERR_ARG, "view range exceeds the plan's views");
  return CTR_OK;
}

static int ensure_buf(float** p, size_t* have, size_t need, const char* what) {
  if (*have >= need) return CTR_OK;
  (void)hipFree(*p);
  *p = nullptr;
  *have = 0;
  if (dev_malloc((void**)p, need * sizeof(float)) != hipSuccess)
    return fail(CTR_ERR_NOMEM, std::string("cannot allocate ") + what);
  *have = need;
  return CTR_OK;
}

// Forward projection with the plan's method (fused residual when b != NULL); T is the
// storage type of vol and out.
template <typename TV, typename TO = TV>
static int fp_any(Plan& P, const TV* vol, TO* out, int vb, int vs, int vc, hipStream_t st,
                  const float* b, const float* invRow, double* partials) {
  if (P.method == CTR_METHOD_PSEUDO_DD || P.method == CTR_METHOD_ZLINE) {
    if constexpr (std::is_same<TV, float>::value && std::is_same<TO, float>::value) {
      if (b) return fail(CTR_ERR_UNSUPPORTED, "the pseudo-DD projector has no fused residual");
      return launch_fp_pdd(P, vol, out, vb, vs, vc, st);
    } else {
      return fail(CTR_ERR_UNSUPPORTED, "the pseudo-DD projector takes FP32 operands");
    }
  }
  if (P.method == CTR_METHOD_SAT) {
    int rc = ensure_buf(&P.sat_vol, &P.sat_vol_n, 2 * sat_volume_floats(P.d),
                        "volume integral images");
    if (rc) return rc;
    if ((rc = launch_volume_sat<TV>(P, vol, P.sat_vol, st))) return rc;
    return launch_fp_sat<TO>(P, P.sat_vol, out, vb, vs, vc, st, b, invRow, partials);
  }
  return launch_fp<TV, TO>(P, vol, out, vb, vs, vc, st, b, invRow, partials);
}

// Back projection (T: storage type of prj; vol FP32); shadow as in launch_bp.
template <typename T>
static int bp_any(Plan& P, const T* prj, float* vol, int vb, int vs, int vc, int kind,
                  hipStream_t st, const float* invCol, int accumulate,
                  uint16_t* shadow = nullptr) {
  if (P.method == CTR_METHOD_PSEUDO_DD || P.method == CTR_METHOD_ZLINE) {
    if (kind != CTR_BP_REFERENCE)
      return fail(CTR_ERR_UNSUPPORTED, "the pseudo-DD / z-line models have no matched adjoint");
    if (accumulate)
      return fail(CTR_ERR_UNSUPPORTED, "the pseudo-DD / z-line back projectors have no update form");
    if constexpr (std::is_same<T, float>::value) {
      if (P.method == CTR_METHOD_PSEUDO_DD) return launch_bp_pdd(P, prj, vol, vb, vs, vc, st);
      const size_t need = zline_prefix_doubles(P.d, vc);
      if (P.zl_cpre_n < need) {
        (void)hipFree(P.zl_cpre);
        P.zl_cpre = nullptr;
        P.zl_cpre_n = 0;
        if (dev_malloc((void**)&P.zl_cpre, std::max<size_t>(need, 1) * sizeof(double)) !=
            hipSuccess)
          return fail(CTR_ERR_NOMEM, "cannot allocate the z-line row prefix");
        P.zl_cpre_n = need;
      }
      return launch_bp_zline(P, prj, P.zl_cpre, vol, vb, vs, vc, st);
    } else {
      return fail(CTR_ERR_UNSUPPORTED, "the pseudo-DD projector takes FP32 operands");
    }
  }
  if (kind == CTR_BP_REFERENCE && P.method == CTR_METHOD_SAT) {
    int rc = ensure_buf(&P.sat_sino, &P.sat_sino_n, 2 * sat_sino_floats(P.d, vc),
                        "sinogram integral images");
    if (rc) return rc;
    return launch_bp_sat<T>(P, prj, P.sat_sino, vol, vb, vs, vc, st, invCol, accumulate, shadow);
  }
  if (kind == CTR_BP_MATCHED) {
    int rc = ensure_scratch_prj(P, vc);
    if (rc) return rc;
    if ((rc = launch_scale_matched<T>(P, prj, P.scratch_prj, vb, vs, vc, st))) return rc;
    return launch_bp<float>(P, P.scratch_prj, vol, vb, vs, vc, kind, st, invCol, accumulate,
                            shadow);
  }
  return launch_bp<T>(P, prj, vol, vb, vs, vc, kind, st, invCol, accumulate, shadow);
}

// The OS-SART and CG drivers run the reference's mode-0 model (os_sart.m:16-27 and
// conjgrad.m call the branchless projectors only).
static int check_solver_method(const Plan& P) {
  if (P.method == CTR_METHOD_PSEUDO_DD || P.method == CTR_METHOD_ZLINE)
    return fail(CTR_ERR_UNSUPPORTED,
                "the OS-SART / CG drivers use the branchless model (method DIRECT or SAT)");
  return CTR_OK;
}

// Owns a set of HIP events.
struct EventPool {
  std::vector<hipEvent_t> e;
  ~EventPool() {
    for (auto x : e) (void)hipEventDestroy(x);
  }
  int init(int n) {
    e.resize(n, nullptr);
    for (auto& x : e) CTR_HIP(hipEventCreate(&x));
    return CTR_OK;
  }
};

// RAII device buffer used by the drivers.
struct DBuf {
  void* p = nullptr;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    if (dev_malloc(&p, bytes) != hipSuccess) {
      p = nullptr;
      return fail(CTR_ERR_NOMEM, "device allocation failed");
    }
    return CTR_OK;
  }
  float* f() const { return (float*)p; }
  double* d() const { return (double*)p; }
  uint16_t* h() const { return (uint16_t*)p; }
};

int residual_run(Plan& P, const float* d_x, const float* d_prj, float* prj, int chunk,
                 double* partials, double* sums, hipStream_t st, double* out) {
  const int PN = P.d.PN;
  const int nch = (PN + chunk - 1) / chunk;
  double h[kResidualSums];
  double rr = 0, bb = 0;
  int rc, pending = 0;
  for (int c = 0; c < nch; ++c) {
    const int v0 = c * chunk, vc = std::min(chunk, PN - v0);
    if ((rc = fp_any(P, d_x, prj, v0, 1, vc, st, d_prj, nullptr, partials))) return rc;
    if ((rc = launch_reduce_pairs(partials, fp_partial_blocks(P, vc), sums + 2 * pending, st)))
      return rc;
    // the chunks' sums are read back kResidualSums / 2 at a time and added in chunk order
    if (++pending == kResidualSums / 2 || c == nch - 1) {
      CTR_HIP(hipMemcpyAsync(h, sums, 2 * pending * sizeof(double), hipMemcpyDeviceToHost, st));
      CTR_HIP(hipStreamSynchronize(st));
      for (int i = 0; i < pending; ++i) {
        rr += h[2 * i];
        bb += h[2 * i + 1];
      }
      pending = 0;
    }
  }
  *out = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
  return CTR_OK;
}

}  // namespace ctr

using namespace ctr;


// An OS-SART session: the subset row / column sums of one measured sinogram, kept so
// that sweeps can be run in several calls (os_sart.m:14-18 computes them once).
struct ctr_sart {
  Plan* P = nullptr;
  const float* prj = nullptr;
  int osNum = 1, kind = CTR_BP_REFERENCE, maxV = 0;
  bool bf = false;      // BF16 shadow of the iterate read by the forward projector
  bool sino16 = false;  // BF16 residual sinogram between the projectors
  double setup_ms = 0.0;
  DBuf rowInv, colInv, tp, xs, sums, parts;
  std::vector<size_t> rowOff;
  // one sweep captured as a HIP graph for the iterate graph_x (replayed per sweep)
  bool use_graph = true;
  float* graph_x = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double capture_ms = 0.0;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graph_x = nullptr;
  }
  ~ctr_sart() { drop_graph(); }
};

namespace {

Plan* get(ctr_plan* p) { return p ? p->impl : nullptr; }

int set_device(int device) {
  int n = 0;
  CTR_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return fail(CTR_ERR_ARG, "device " + std::to_string(device) + " not present (" +
                                 std::to_string(n) + " visible)");
  CTR_HIP(hipSetDevice(device));
  return CTR_OK;
}

ctr_geometry make_geometry(float x0, float y0, float z0, int DNU, int DNV, const float* xds,
                           const float* yds, const float* zds, float cx, float cy, float cz,
                           const float* angs, const float* zpos, int PN, int XN, int YN, int ZN,
                           float dx, float dz, const unsigned char* mask) {
  ctr_geometry g;
  g.x0 = x0; g.y0 = y0; g.z0 = z0; g.DNU = DNU; g.DNV = DNV;
  g.xds = xds; g.yds = yds; g.zds = zds;
  g.imgXCenter = cx; g.imgYCenter = cy; g.imgZCenter = cz;
  g.angles = angs; g.zshifts = zpos; g.PN = PN;
  g.XN = XN; g.YN = YN; g.ZN = ZN; g.dx = dx; g.dz = dz; g.mask = mask;
  return g;
}

// prjMode dispatch of DD3Proj_gpu / DD3Back_gpu (DD3_GPU.cu:1144-1183, 1778-1815):
// what each value selects, every value outside 1..3 being mode 0.
enum class Model { Branchless, PseudoDD, Fp64Sat, ZLine };
Model fp_model(int mode) {
  if (mode == 1 || mode == 3) return Model::PseudoDD;
  if (mode == 2) return Model::Fp64Sat;
  return Model::Branchless;
}
Model bp_model(int mode) {
  if (mode == 1 || mode == 2) return Model::PseudoDD;
  if (mode == 3) return Model::ZLine;
  return Model::Branchless;
}

struct PlanGuard {
  ctr_plan* p = nullptr;
  ~PlanGuard() {
    if (p) ctr_plan_destroy(p);
  }
};

}  // namespace

extern "C" {

const char* ctr_last_error(void) { return last_error_cstr(); }

const char* ctr_version(void) { return "ctrecon_amd 0.1 (gfx950)"; }

// A plan on `device` with tables built at offset `half` (build_tables) and its stream.
static int make_plan(const ctr_geometry* geom, int device, double half, ctr_plan** out) {
  if (!out) return fail(CTR_ERR_ARG, "out is NULL");
  *out = nullptr;
  int rc = validate_geometry(geom);
  if (rc) return rc;
  if ((rc = set_device(device))) return rc;
  Plan* P = new Plan();
  P->device = device;
  if ((rc = build_tables(geom, *P, half))) {
    free_plan(P);
    return rc;
  }
  if (hipStreamCreate(&P->stream) != hipSuccess) {
    P->stream = nullptr;
    free_plan(P);
    return fail(CTR_ERR_HIP, "cannot create the plan's stream");
  }
  *out = new ctr_plan{P};
  set_error("");
  return CTR_OK;
}

int ctr_plan_create(const ctr_geometry* geom, int device, ctr_plan** out) {
  return make_plan(geom, device, 0.0, out);
}

int ctr_plan_destroy(ctr_plan* plan) {
  if (!plan) return CTR_OK;
  if (plan->impl) {
    (void)hipSetDevice(plan->impl->device);
    free_plan(plan->impl);
  }
  delete plan;
  return CTR_OK;
}

int ctr_plan_set_method(ctr_plan* plan, int method) {
  Plan* P = get(plan);
  if (!P) return fail(CTR_ERR_ARG, "NULL plan");
  if (method != CTR_METHOD_DIRECT && method != CTR_METHOD_SAT &&
      method != CTR_METHOD_PSEUDO_DD && method != CTR_METHOD_ZLINE)
    return fail(CTR_ERR_ARG, "method must be CTR_METHOD_DIRECT, CTR_METHOD_SAT, "
                             "CTR_METHOD_PSEUDO_DD or CTR_METHOD_ZLINE");
  P->method = method;
  return CTR_OK;
}

int ctr_plan_fp(ctr_plan* plan, const float* d_vol, float* d_prj, int view_begin,
                int view_stride, int view_count, void* stream) {
  Plan* P = get(plan);
  if (!P || !d_vol || !d_prj) return fail(CTR_ERR_ARG, "NULL plan or buffer");
  int rc = check_views(*P, view_begin, view_stride, view_count);
  if (rc) return rc;
  CTR_HIP(hipSetDevice(P->device));
  return fp_any(*P, d_vol, d_prj, view_begin, view_stride, view_count, (hipStream_t)stream,
                   nullptr, nullptr, nullptr);
}

int ctr_plan_bp(ctr_plan* plan, const float* d_prj, float* d_vol, int view_begin,
                int view_stride, int view_count, int bp_kind, void* stream) {
  Plan* P = get(plan);
  if (!P || !d_vol || !d_prj) return fail(CTR_ERR_ARG, "NULL plan or buffer");
  if (bp_kind != CTR_BP_REFERENCE && bp_kind != CTR_BP_MATCHED)
    return fail(CTR_ERR_ARG, "bp_kind must be CTR_BP_REFERENCE or CTR_BP_MATCHED");
  int rc = check_views(*P, view_begin, view_stride, view_count);
  if (rc) return rc;
  CTR_HIP(hipSetDevice(P->device));
  return bp_any(*P, d_prj, d_vol, view_begin, view_stride, view_count, bp_kind,
                (hipStream_t)stream, nullptr, 0);
}

int ctr_plan_set_storage(ctr_plan* plan, int storage) {
  Plan* P = get(plan);
  if (!P) return fail(CTR_ERR_ARG, "NULL plan");
  if (storage != CTR_STORAGE_F32 && storage != CTR_STORAGE_BF16 &&
      storage != CTR_STORAGE_BF16_SINO)
    return fail(CTR_ERR_ARG,
                "storage must be CTR_STORAGE_F32, CTR_STORAGE_BF16 or CTR_STORAGE_BF16_SINO");
  P->storage = storage;
  return CTR_OK;
}

int ctr_plan_mask_volume(ctr_plan* plan, float* d_vol, void* stream) {
  Plan* P = get(plan);
  if (!P || !d_vol) return fail(CTR_ERR_ARG, "NULL plan or buffer");
  CTR_HIP(hipSetDevice(P->device));
  return launch_mask_volume(*P, d_vol, (hipStream_t)stream);
}

int ctr_plan_fp_bf16(ctr_plan* plan, const unsigned short* d_vol, unsigned short* d_prj,
                     int view_begin, int view_stride, int view_count, void* stream) {
  Plan* P = get(plan);
  if (!P || !d_vol || !d_prj) return fail(CTR_ERR_ARG, "NULL plan or buffer");
  int rc = check_views(*P, view_begin, view_stride, view_count);
  if (rc) return rc;
  CTR_HIP(hipSetDevice(P->device));
  return fp_any<uint16_t>(*P, d_vol, d_prj, view_begin, view_stride, view_count,
                          (hipStream_t)stream, nullptr, nullptr, nullptr);
}

int ctr_plan_bp_bf16(ctr_plan* plan, const unsigned short* d_prj, float* d_vol, int view_begin,
                     int view_stride, int view_count, int bp_kind, void* stream) {
  Plan* P = get(plan);
  if (!P || !d_vol || !d_prj) return fail(CTR_ERR_ARG, "NULL plan or buffer");
  if (bp_kind != CTR_BP_REFERENCE && bp_kind != CTR_BP_MATCHED)
    return fail(CTR_ERR_ARG, "bp_kind must be CTR_BP_REFERENCE or CTR_BP_MATCHED");
  int rc = check_views(*P, view_begin, view_stride, view_count);
  if (rc) return rc;
  CTR_HIP(hipSetDevice(P->device));
  return bp_any<uint16_t>(*P, d_prj, d_vol, view_begin, view_stride, view_count, bp_kind,
                          (hipStream_t)stream, nullptr, 0);
}

int ctr_f32_to_bf16(unsigned short* d_dst, const float* d_src, unsigned long long n,
                    void* stream) {
  if (!d_dst || !d_src) return fail(CTR_ERR_ARG, "NULL pointer");
  return launch_to_bf16(d_dst, d_src, n, (hipStream_t)stream);
}

int ctr_bf16_to_f32(float* d_dst, const unsigned short* d_src, unsigned long long n,
                    void* stream) {
  if (!d_dst || !d_src) return fail(CTR_ERR_ARG, "NULL pointer");
  return launch_from_bf16(d_dst, d_src, n, (hipStream_t)stream);
}

int ctr_residual(ctr_plan* plan, const float* d_x, const float* d_prj, double* out) {
  Plan* P = get(plan);
  if (!P || !d_x || !d_prj || !out) return fail(CTR_ERR_ARG, "NULL argument");
  if (int rc0 = check_solver_method(*P)) return rc0;
  CTR_HIP(hipSetDevice(P->device));
  // chunk the views so the scratch stays modest
  const int chunk = std::min(P->d.PN, 256);
  int rc = ensure_scratch_prj(*P, chunk);
  if (rc) return rc;
  if ((rc = ensure_partials(*P, 2 * (size_t)fp_partial_blocks(*P, chunk) + 2))) return rc;
  size_t have = P->sums2 ? kResidualSums * sizeof(double) : 0;
  if ((rc = ensure_bytes((void**)&P->sums2, &have, kResidualSums * sizeof(double),
                         "residual sums")))
    return rc;
  return residual_run(*P, d_x, d_prj, P->scratch_prj, chunk, P->partials, P->sums2, P->stream,
                      out);
}

int ctr_os_sart(ctr_plan* plan, float* d_x, const float* d_prj, int osNum, int iterNum,
                int bp_kind, double* resid) {
  return ctr_os_sart_run(plan, d_x, d_prj, osNum, iterNum, bp_kind, 0.0, resid, nullptr, nullptr);
}

int ctr_os_sart_run(ctr_plan* plan, float* d_x, const float* d_prj, int osNum, int maxIter,
                    int bp_kind, double stop_resid, double* resid, double* sweep_ms,
                    ctr_sart_stats* stats) {
  ctr_sart* S = nullptr;
  int rc = ctr_sart_create(plan, d_prj, osNum, bp_kind, &S);
  if (rc) return rc;
  rc = ctr_sart_sweeps(S, d_x, maxIter, stop_resid, resid, sweep_ms, stats);
  ctr_sart_destroy(S);
  return rc;
}

int ctr_sart_create(ctr_plan* plan, const float* d_prj, int osNum, int bp_kind, ctr_sart** out) {
  if (!out) return fail(CTR_ERR_ARG, "out is NULL");
  *out = nullptr;
  Plan* P = get(plan);
  if (!P || !d_prj) return fail(CTR_ERR_ARG, "NULL argument");
  if (osNum < 1 || osNum > P->d.PN) return fail(CTR_ERR_ARG, "osNum must be in [1, PN]");
  if (bp_kind != CTR_BP_REFERENCE && bp_kind != CTR_BP_MATCHED)
    return fail(CTR_ERR_ARG, "bad bp_kind");
  if (int rc0 = check_solver_method(*P)) return rc0;
  CTR_HIP(hipSetDevice(P->device));
  std::unique_ptr<ctr_sart> S(new ctr_sart());
  S->P = P;
  S->prj = d_prj;
  S->osNum = osNum;
  S->kind = bp_kind;
  S->bf = P->storage == CTR_STORAGE_BF16;
  S->sino16 = P->storage != CTR_STORAGE_F32;
  const Dims& d = P->d;
  const size_t NV = (size_t)d.XN * d.YN * d.ZN, frame = (size_t)d.DNU * d.DNV;
  S->maxV = (d.PN + osNum - 1) / osNum;
  int rc;
  if ((rc = S->rowInv.alloc(frame * d.PN * sizeof(float)))) return rc;  // subset-compact order
  if ((rc = S->colInv.alloc(NV * osNum * sizeof(float)))) return rc;
  if ((rc = S->tp.alloc(frame * S->maxV * (S->sino16 ? sizeof(uint16_t) : sizeof(float)))))
    return rc;
  if (S->bf && (rc = S->xs.alloc(NV * sizeof(uint16_t)))) return rc;
  if ((rc = S->sums.alloc(2 * sizeof(double) * osNum))) return rc;
  if ((rc = S->parts.alloc((2 * (size_t)fp_partial_blocks(*P, S->maxV) + 2) * sizeof(double))))
    return rc;
  hipStream_t st = P->stream;
  EventPool ev;
  if ((rc = ev.init(2))) return rc;
  CTR_HIP(hipEventRecord(ev.e[0], st));
  // RowSum{s} = fp(ones), ColSum{s} = bp(ones) (os_sart.m:14-18).  The branchless
  // direct method evaluates them without a volume / sinogram of ones (launch_fp_ones,
  // launch_bp_ones); the other methods, the matched adjoint and columns over the LDS
  // limit project an array of ones.
  const bool ones_fp = P->method == CTR_METHOD_DIRECT;
  const bool ones_bp = ones_fp && bp_kind == CTR_BP_REFERENCE && d.ZN <= kWinMaxZ;
  DBuf ones;
  if (!ones_fp || !ones_bp) {
    if ((rc = ones.alloc(std::max(NV, frame * S->maxV) * sizeof(float)))) return rc;
    if ((rc = launch_fill(ones.f(), std::max(NV, frame * S->maxV), 1.0f, st))) return rc;
  }
  S->rowOff.resize(osNum);
  size_t off = 0;
  for (int s = 0; s < osNum; ++s) {
    const int vc = (d.PN - s + osNum - 1) / osNum;
    S->rowOff[s] = off;
    float* rI = S->rowInv.f() + off * frame;
    rc = ones_fp ? launch_fp_ones(*P, rI, s, osNum, vc, st)
                 : fp_any<float>(*P, ones.f(), rI, s, osNum, vc, st, nullptr, nullptr, nullptr);
    if (rc) return rc;
    if ((rc = launch_inv_guard(rI, rI, frame * vc, st))) return rc;
    float* cI = S->colInv.f() + NV * s;
    rc = ones_bp ? launch_bp_ones(*P, cI, s, osNum, vc, st)
                 : bp_any<float>(*P, ones.f(), cI, s, osNum, vc, bp_kind, st, nullptr, 0);
    if (rc) return rc;
    if ((rc = launch_inv_guard(cI, cI, NV, st))) return rc;
    off += vc;
  }
  CTR_HIP(hipEventRecord(ev.e[1], st));
  CTR_HIP(hipEventSynchronize(ev.e[1]));
  float ms = 0;
  CTR_HIP(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
  S->setup_ms = ms;
  *out = S.release();
  return CTR_OK;
}

int ctr_sart_destroy(ctr_sart* S) {
  if (S) {
    (void)hipSetDevice(S->P->device);
    (void)hipStreamSynchronize(S->P->stream);
    delete S;
  }
  return CTR_OK;
}

double ctr_sart_setup_ms(const ctr_sart* S) { return S ? S->setup_ms : 0.0; }

// The launches of one sweep: per subset the fused forward projection (residual sinogram
// and its sums), the reduction of the sums, and the back-projection update.  kev (may be
// NULL): four events per subset around the two projector launches.
static int sweep_launches(ctr_sart* S, float* d_x, hipStream_t st, hipEvent_t* kev) {
  Plan* P = S->P;
  const Dims& d = P->d;
  const int osNum = S->osNum;
  const size_t NV = (size_t)d.XN * d.YN * d.ZN, frame = (size_t)d.DNU * d.DNV;
  int rc;
  for (int s = 0; s < osNum; ++s) {
    const int vc = (d.PN - s + osNum - 1) / osNum;
    if (kev) CTR_HIP(hipEventRecord(kev[4 * s + 0], st));
    // tp = (prj - A_s x) ./ RowSum (os_sart.m:24,30-31)
    const float* rI = S->rowInv.f() + S->rowOff[s] * frame;
    if (S->bf)
      rc = fp_any<uint16_t, uint16_t>(*P, S->xs.h(), S->tp.h(), s, osNum, vc, st, S->prj, rI,
                                      S->parts.d());
    else if (S->sino16)
      rc = fp_any<float, uint16_t>(*P, d_x, S->tp.h(), s, osNum, vc, st, S->prj, rI,
                                   S->parts.d());
    else
      rc = fp_any<float, float>(*P, d_x, S->tp.f(), s, osNum, vc, st, S->prj, rI, S->parts.d());
    if (rc) return rc;
    if (kev) CTR_HIP(hipEventRecord(kev[4 * s + 1], st));
    if ((rc = launch_reduce_pairs(S->parts.d(), fp_partial_blocks(*P, vc), S->sums.d() + 2 * s,
                                  st)))
      return rc;
    if (kev) CTR_HIP(hipEventRecord(kev[4 * s + 2], st));
    // x += B_s(tp) ./ ColSum (os_sart.m:32,37-40)
    const float* cI = S->colInv.f() + NV * s;
    rc = S->sino16 ? bp_any<uint16_t>(*P, S->tp.h(), d_x, s, osNum, vc, S->kind, st, cI, 1,
                                      S->bf ? S->xs.h() : nullptr)
                   : bp_any<float>(*P, S->tp.f(), d_x, s, osNum, vc, S->kind, st, cI, 1);
    if (rc) return rc;
    if (kev) CTR_HIP(hipEventRecord(kev[4 * s + 3], st));
  }
  return CTR_OK;
}

// A sweep can be replayed from a graph when every buffer its launches touch belongs to
// the session or the caller: the direct method with the reference back projector (the
// integral-image method and the matched adjoint use plan scratch that later calls may
// reallocate).
static bool graph_capturable(const ctr_sart* S) {
  return S->P->method == CTR_METHOD_DIRECT && S->kind == CTR_BP_REFERENCE;
}

static int capture_sweep(ctr_sart* S, float* d_x, hipStream_t st) {
  S->drop_graph();
  CTR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = sweep_launches(S, d_x, st, nullptr);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess || !g)
    return fail(CTR_ERR_HIP, std::string("sweep graph capture: ") + hipGetErrorString(e));
  S->graph = g;
  if (hipGraphInstantiate(&S->exec, S->graph, nullptr, nullptr, 0) != hipSuccess) {
    S->exec = nullptr;
    S->drop_graph();
    return fail(CTR_ERR_HIP, "sweep graph instantiation failed");
  }
  S->graph_x = d_x;
  return CTR_OK;
}

int ctr_sart_sweeps(ctr_sart* S, float* d_x, int maxIter, double stop_resid, double* resid,
                    double* sweep_ms, ctr_sart_stats* stats) {
  if (!S || !d_x) return fail(CTR_ERR_ARG, "NULL argument");
  if (maxIter < 0) return fail(CTR_ERR_ARG, "iterNum must be >= 0");
  Plan* P = S->P;
  CTR_HIP(hipSetDevice(P->device));
  const Dims& d = P->d;
  const int osNum = S->osNum;
  const size_t NV = (size_t)d.XN * d.YN * d.ZN;
  int rc;
  hipStream_t st = P->stream;
  const bool timing = stats != nullptr || sweep_ms != nullptr;
  // per-launch timing (stats) needs the events between the launches: plain launches then
  const bool graph = S->use_graph && stats == nullptr && graph_capturable(S) && maxIter > 0;
  // events: [0] start, then one per sweep end; per-launch pairs for the kernel totals
  EventPool ev;
  if (timing && (rc = ev.init(1 + maxIter + 4 * osNum))) return rc;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->setup_ms = S->setup_ms;
  }
  if (graph && (S->exec == nullptr || S->graph_x != d_x)) {
    const auto t0 = std::chrono::steady_clock::now();
    if ((rc = capture_sweep(S, d_x, st))) return rc;
    S->capture_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  if (timing) CTR_HIP(hipEventRecord(ev.e[0], st));
  // the caller may have changed x since the last call: refresh its BF16 shadow
  if (S->bf && (rc = launch_to_bf16_shadow(S->xs.h(), d_x, NV, st))) return rc;
  hipEvent_t* kev = stats ? ev.e.data() + 1 + maxIter : nullptr;
  std::vector<double> h(2 * osNum);
  int done = 0;
  for (int it = 0; it < maxIter; ++it) {
    if (graph) {
      CTR_HIP(hipGraphLaunch(S->exec, st));
    } else if ((rc = sweep_launches(S, d_x, st, kev))) {
      return rc;
    }
    if (timing) CTR_HIP(hipEventRecord(ev.e[1 + it], st));
    CTR_HIP(hipMemcpyAsync(h.data(), S->sums.p, h.size() * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    CTR_HIP(hipStreamSynchronize(st));
    double rr = 0, bb = 0;
    for (int s = 0; s < osNum; ++s) {
      rr += h[2 * s];
      bb += h[2 * s + 1];
    }
    const double r = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    if (resid) resid[it] = r;
    if (stats) {
      for (int s = 0; s < osNum; ++s) {
        float a = 0, b = 0;
        CTR_HIP(hipEventElapsedTime(&a, kev[4 * s + 0], kev[4 * s + 1]));
        CTR_HIP(hipEventElapsedTime(&b, kev[4 * s + 2], kev[4 * s + 3]));
        stats->fp_ms += a;
        stats->bp_ms += b;
      }
      stats->fp_launches += osNum;
      stats->bp_launches += osNum;
    }
    if (sweep_ms) {
      float ms = 0;
      CTR_HIP(hipEventElapsedTime(&ms, ev.e[it], ev.e[1 + it]));
      sweep_ms[it] = ms;
    }
    done = it + 1;
    if (stop_resid > 0 && r <= stop_resid) break;
  }
  CTR_HIP(hipStreamSynchronize(st));
  if (stats) stats->sweeps = done;
  return CTR_OK;
}

int ctr_sart_set_graph(ctr_sart* S, int enable) {
  if (!S) return fail(CTR_ERR_ARG, "NULL session");
  S->use_graph = enable != 0;
  if (!S->use_graph) {
    (void)hipSetDevice(S->P->device);
    S->drop_graph();
  }
  return CTR_OK;
}

int ctr_sart_graph_info(const ctr_sart* S, int* captured, double* capture_ms) {
  if (!S) return fail(CTR_ERR_ARG, "NULL session");
  if (captured) *captured = S->exec != nullptr;
  if (capture_ms) *capture_ms = S->capture_ms;
  return CTR_OK;
}

int ctr_cone_voxels(ctr_plan* plan, int view_begin, int view_stride, int view_count,
                    long long* out) {
  Plan* P = get(plan);
  if (!P || !out) return fail(CTR_ERR_ARG, "NULL argument");
  int rc = check_views(*P, view_begin, view_stride, view_count);
  if (rc) return rc;
  const HostGeo& h = P->hg;
  const Dims& d = P->d;
  for (int k = 0; k < view_count; ++k) {
    const int p = view_begin + k * view_stride;
    const double sx = P->h_views[p].sx, sy = P->h_views[p].sy, sz = h.sz[p];
    long long n = 0;
    for (int y = 0; y < d.YN; ++y)
      for (int x = 0; x < d.XN; ++x) {
        if (!h.mask[(size_t)y * d.XN + x]) continue;
        const double Px = (x - h.ocx) * h.dx - sx, Py = (y - h.ocy) * h.dx - sy;
        const double l = std::sqrt(Px * Px + Py * Py) / h.S2D;
        const double zlo = sz + h.bz0 * l, zhi = sz + h.bzN * l;
        int k0 = (int)std::floor(zlo / h.dz + h.ocz - 0.5) + 1;
        int k1 = (int)std::ceil(zhi / h.dz + h.ocz + 0.5) - 1;
        k0 = std::max(k0, 0);
        k1 = std::min(k1, d.ZN - 1);
        if (k1 >= k0) n += k1 - k0 + 1;
      }
    out[k] = n;
  }
  return CTR_OK;
}

int ctr_conjgrad(ctr_plan* plan, float* d_x, const float* d_prj, int iterNum, int bp_kind,
                 double* resid) {
  Plan* P = get(plan);
  if (!P || !d_x || !d_prj) return fail(CTR_ERR_ARG, "NULL argument");
  if (iterNum < 0) return fail(CTR_ERR_ARG, "iterNum must be >= 0");
  if (bp_kind != CTR_BP_REFERENCE && bp_kind != CTR_BP_MATCHED)
    return fail(CTR_ERR_ARG, "bad bp_kind");
  if (int rc0 = check_solver_method(*P)) return rc0;
  CTR_HIP(hipSetDevice(P->device));
  const Dims& d = P->d;
  const size_t NV = (size_t)d.XN * d.YN * d.ZN, NS = (size_t)d.DNU * d.DNV * d.PN;
  int rc;
  // BF16 storage: A reads a BF16 copy of its input and passes a BF16 sinogram between
  // the projectors; the CG vectors stay FP32 and the dots / scalars FP64
  const bool bf = P->storage == CTR_STORAGE_BF16;
  const bool sino16 = P->storage != CTR_STORAGE_F32;
  // the CG vectors live in the plan: the first call allocates them, later calls reuse them
  const size_t need[5] = {NV * sizeof(float), NV * sizeof(float), NV * sizeof(float),
                          NS * (sino16 ? sizeof(uint16_t) : sizeof(float)),
                          bf ? NV * sizeof(uint16_t) : 0};
  const char* what[5] = {"CG residual", "CG direction", "CG operator output",
                         "CG sinogram", "CG BF16 input"};
  for (int i = 0; i < 5; ++i)
    if (need[i] && (rc = ensure_bytes(&P->cg[i], &P->cg_bytes[i], need[i], what[i]))) return rc;
  {
    size_t have = P->sums2 ? kResidualSums * sizeof(double) : 0;
    if ((rc = ensure_bytes((void**)&P->sums2, &have, kResidualSums * sizeof(double),
                           "CG dot products")))
      return rc;
  }
  const int nparts = 1024;
  if ((rc = ensure_partials(*P, 2 * (size_t)nparts + 2))) return rc;
  struct V {
    void* p;
    float* f() const { return (float*)p; }
    uint16_t* h() const { return (uint16_t*)p; }
  };
  const V r{P->cg[0]}, pv{P->cg[1]}, t{P->cg[2]}, sino{P->cg[3]}, in16{P->cg[4]};
  double* dots = P->sums2;
  hipStream_t st = P->stream;
  auto A = [&](const float* in, float* out) -> int {  // Back(Proj(.)) (conjgrad.m:4-6)
    if (bf) {
      int e = launch_to_bf16_shadow(in16.h(), in, NV, st);
      if (e) return e;
      if ((e = fp_any<uint16_t>(*P, in16.h(), sino.h(), 0, 1, d.PN, st, nullptr, nullptr,
                                nullptr)))
        return e;
      return bp_any<uint16_t>(*P, sino.h(), out, 0, 1, d.PN, bp_kind, st, nullptr, 0);
    }
    if (sino16) {
      int e = fp_any<float, uint16_t>(*P, in, sino.h(), 0, 1, d.PN, st, nullptr, nullptr,
                                      nullptr);
      if (e) return e;
      return bp_any<uint16_t>(*P, sino.h(), out, 0, 1, d.PN, bp_kind, st, nullptr, 0);
    }
    int e = fp_any<float, float>(*P, in, sino.f(), 0, 1, d.PN, st, nullptr, nullptr, nullptr);
    if (e) return e;
    return bp_any<float>(*P, sino.f(), out, 0, 1, d.PN, bp_kind, st, nullptr, 0);
  };
  auto dot = [&](const float* a, const float* b, double* outv) -> int {
    int e = launch_dot(a, b, NV, P->partials, nparts, dots, st);
    if (e) return e;
    double h[2];
    if (hipMemcpyAsync(h, dots, sizeof(h), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(CTR_ERR_HIP, "dot readback failed");
    *outv = h[0];
    return CTR_OK;
  };
  // b = Back(prj); r0 = b - A(x0); p0 = r0 (conjgrad.m:7-9)
  if ((rc = bp_any(*P, d_prj, r.f(), 0, 1, d.PN, bp_kind, st, nullptr, 0))) return rc;
  if ((rc = A(d_x, t.f()))) return rc;
  if ((rc = launch_sub(r.f(), r.f(), t.f(), NV, st))) return rc;
  CTR_HIP(hipMemcpyAsync(pv.p, r.p, NV * sizeof(float), hipMemcpyDeviceToDevice, st));
  double rr;
  if ((rc = dot(r.f(), r.f(), &rr))) return rc;
  for (int it = 0; it < iterNum; ++it) {
    if ((rc = A(pv.f(), t.f()))) return rc;  // conjgrad.m:14
    double pAp;
    if ((rc = dot(pv.f(), t.f(), &pAp))) return rc;
    const double alpha = rr / pAp;  // conjgrad.m:18
    if ((rc = launch_axpy(d_x, pv.f(), alpha, NV, st))) return rc;    // :19
    if ((rc = launch_axpy(r.f(), t.f(), -alpha, NV, st))) return rc;  // :20
    double r1;
    if ((rc = dot(r.f(), r.f(), &r1))) return rc;
    const double beta = r1 / rr;  // :24
    if ((rc = launch_xpby(pv.f(), r.f(), beta, NV, st))) return rc;  // :25
    rr = r1;
    if (resid) resid[it] = std::sqrt(r1);
  }
  CTR_HIP(hipStreamSynchronize(st));
  return CTR_OK;
}

int ctr_dd3_proj(float x0, float y0, float z0, int DNU, int DNV, const float* xds,
                 const float* yds, const float* zds, float imgXCenter, float imgYCenter,
                 float imgZCenter, const float* hangs, const float* hzPos, int PN, int XN,
                 int YN, int ZN, const float* hvol, float* hprj, float dx, float dz,
                 const unsigned char* mask, int gpunum, int prjMode) {
  if (!hvol || !hprj) return fail(CTR_ERR_ARG, "NULL volume or projection buffer");
  const Model m = fp_model(prjMode);
  ctr_geometry g = make_geometry(x0, y0, z0, DNU, DNV, xds, yds, zds, imgXCenter, imgYCenter,
                                 imgZCenter, hangs, hzPos, PN, XN, YN, ZN, dx, dz, mask);
  PlanGuard pg;
  int rc;
  // prjMode 2: the branchless footprint model, sampled half a cell off (build_tables)
  if ((rc = make_plan(&g, gpunum, m == Model::Fp64Sat ? 0.5 : 0.0, &pg.p))) return rc;
  if (m == Model::PseudoDD && (rc = ctr_plan_set_method(pg.p, CTR_METHOD_PSEUDO_DD))) return rc;
  const size_t NV = (size_t)XN * YN * ZN, NS = (size_t)DNU * DNV * PN;
  DBuf dv, dp;
  if ((rc = dv.alloc(NV * sizeof(float)))) return rc;
  if ((rc = dp.alloc(NS * sizeof(float)))) return rc;
  CTR_HIP(hipMemcpy(dv.p, hvol, NV * sizeof(float), hipMemcpyHostToDevice));
  if ((rc = ctr_plan_fp(pg.p, dv.f(), dp.f(), 0, 1, PN, nullptr))) return rc;
  CTR_HIP(hipMemcpy(hprj, dp.p, NS * sizeof(float), hipMemcpyDeviceToHost));
  return CTR_OK;
}

int ctr_dd3_back(float x0, float y0, float z0, int DNU, int DNV, const float* xds,
                 const float* yds, const float* zds, float imgXCenter, float imgYCenter,
                 float imgZCenter, const float* hangs, const float* hzPos, int PN, int XN,
                 int YN, int ZN, float* hvol, const float* hprj, float dx, float dz,
                 const unsigned char* mask, int gpunum, int squared, int prjMode) {
  (void)squared;  // ignored, as in the reference (DD3_GPU.cu:1393)
  if (!hvol || !hprj) return fail(CTR_ERR_ARG, "NULL volume or projection buffer");
  const bool matched = (prjMode & 16) != 0;
  const Model m = bp_model(prjMode & ~16);
  if (matched && m != Model::Branchless)
    return fail(CTR_ERR_UNSUPPORTED, "the matched adjoint exists for the branchless model only");
  ctr_geometry g = make_geometry(x0, y0, z0, DNU, DNV, xds, yds, zds, imgXCenter, imgYCenter,
                                 imgZCenter, hangs, hzPos, PN, XN, YN, ZN, dx, dz, mask);
  PlanGuard pg;
  int rc;
  if ((rc = ctr_plan_create(&g, gpunum, &pg.p))) return rc;
  if (m == Model::PseudoDD && (rc = ctr_plan_set_method(pg.p, CTR_METHOD_PSEUDO_DD))) return rc;
  if (m == Model::ZLine && (rc = ctr_plan_set_method(pg.p, CTR_METHOD_ZLINE))) return rc;
  const size_t NV = (size_t)XN * YN * ZN, NS = (size_t)DNU * DNV * PN;
  DBuf dv, dp;
  if ((rc = dv.alloc(NV * sizeof(float)))) return rc;
  if ((rc = dp.alloc(NS * sizeof(float)))) return rc;
  CTR_HIP(hipMemcpy(dp.p, hprj, NS * sizeof(float), hipMemcpyHostToDevice));
  if ((rc = ctr_plan_bp(pg.p, dp.f(), dv.f(), 0, 1, PN,
                        matched ? CTR_BP_MATCHED : CTR_BP_REFERENCE, nullptr)))
    return rc;
  CTR_HIP(hipMemcpy(hvol, dv.p, NV * sizeof(float), hipMemcpyDeviceToHost));
  return CTR_OK;
}

int ctr_device_count(int* count) {
  if (!count) return fail(CTR_ERR_ARG, "count is NULL");
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return fail(CTR_ERR_HIP, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
  }
  return CTR_OK;
}

int ctr_malloc(void** d_ptr, unsigned long long bytes, int device) {
  if (!d_ptr) return fail(CTR_ERR_ARG, "d_ptr is NULL");
  int rc = set_device(device);
  if (rc) return rc;
  if (dev_malloc(d_ptr, bytes) != hipSuccess) {
    *d_ptr = nullptr;
    return fail(CTR_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  return CTR_OK;
}

int ctr_free(void* d_ptr) {
  if (d_ptr) CTR_HIP(hipFree(d_ptr));
  return CTR_OK;
}

int ctr_memcpy_h2d(void* d_dst, const void* h_src, unsigned long long bytes) {
  if (!d_dst || !h_src) return fail(CTR_ERR_ARG, "NULL pointer");
  CTR_HIP(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
  return CTR_OK;
}

int ctr_memcpy_d2h(void* h_dst, const void* d_src, unsigned long long bytes) {
  if (!h_dst || !d_src) return fail(CTR_ERR_ARG, "NULL pointer");
  CTR_HIP(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
  return CTR_OK;
}

int ctr_memset(void* d_dst, int value, unsigned long long bytes) {
  if (!d_dst) return fail(CTR_ERR_ARG, "NULL pointer");
  CTR_HIP(hipMemset(d_dst, value, bytes));
  return CTR_OK;
}

int ctr_device_copy(float* d_dst, const float* d_src, unsigned long long n, void* stream) {
  if (!d_dst || !d_src) return fail(CTR_ERR_ARG, "NULL pointer");
  return launch_copy(d_dst, d_src, n, (hipStream_t)stream);
}

int ctr_device_sub(float* d_out, const float* d_a, const float* d_b, unsigned long long n,
                   void* stream) {
  if (!d_out || !d_a || !d_b) return fail(CTR_ERR_ARG, "NULL pointer");
  return launch_sub(d_out, d_a, d_b, n, (hipStream_t)stream);
}

int ctr_device_axpy(float* d_y, const float* d_x, double a, unsigned long long n, void* stream) {
  if (!d_y || !d_x) return fail(CTR_ERR_ARG, "NULL pointer");
  return launch_axpy(d_y, d_x, a, n, (hipStream_t)stream);
}

int ctr_debug_fail_alloc_after(long long n) {
  g_alloc_countdown.store(n < 0 ? -1 : n);
  return CTR_OK;
}

int ctr_synchronize(int device) {
  int rc = set_device(device);
  if (rc) return rc;
  CTR_HIP(hipDeviceSynchronize());
  return CTR_OK;
}

}  // extern "C"
