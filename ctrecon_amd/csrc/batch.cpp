// Batched OS-SART over several helical scans of uneven size (the C ABI ctr_batch_*).
//
// One sweep is osNum subset steps; in step s every scan forward-projects its subset s
// (fused residual), the residual partials of every scan are reduced, and every scan
// back-projects and updates its iterate.  Each of the three is ONE launch over all
// scans: persistent workgroups walk a flat work list whose per-scan ranges and
// operands come from descriptor tables built here once (device memory, one table per
// subset step).  A whole sweep (3 * osNum launches) is captured into a HIP graph and
// replayed; the residual readback between sweeps stays outside the graph.  The
// persistent grid sizes, the forward work mapping and the back projector's z groups
// per lane are chosen by timing the candidates on the batch itself (autotune); the
// choice is kept in a cache file keyed by the device and the batch's geometry, so a
// later batch of the same scans skips the timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "ctr_internal.hpp"

using namespace ctr;

namespace {

struct Mem {
  void* p = nullptr;
  ~Mem() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    if (dev_malloc(&p, bytes) != hipSuccess) {
      p = nullptr;
      return fail(CTR_ERR_NOMEM, "batch: device allocation failed");
    }
    return CTR_OK;
  }
  float* f() const { return (float*)p; }
  double* d() const { return (double*)p; }
  uint16_t* h() const { return (uint16_t*)p; }
};

struct Scan {
  Plan* P = nullptr;
  float* x = nullptr;
  const float* prj = nullptr;
  size_t NV = 0, frame = 0;
  Mem rowInv, colInv, tp, xs, partials;
  std::vector<size_t> rowOff;
};

int vcount(const Plan& P, int s, int osNum) { return (P.d.PN - s + osNum - 1) / osNum; }

}  // namespace

struct ctr_batch {
  int n = 0, osNum = 1, kind = CTR_BP_REFERENCE, device = 0, ncu = 256;
  bool bf = false;
  std::vector<std::unique_ptr<Scan>> scans;
  Mem sums, zeros;  // sums: [n][osNum][2]; zeros: the largest volume (tuning)
  Mem fpItems, redItems;
  std::vector<int> fpTotal;
  // back-projection work lists by class: c = 1..4 z groups per lane, c = 0 the windowed
  // form (one wave per column over all z); the scans of the class, and per subset step
  // their items (and timing copies with a zero column sum)
  std::vector<int> clsScans[5];
  Mem bpItems[5], tuneBpItems[5];
  std::vector<int> bpTotal[5];
  int winZ = 0;    // tallest column of the windowed class
  int winDNV = 0;  // most detector rows of a scan of the windowed class
  bool colw = false;
  // zg: z groups per lane for every scan (1..4), 0 for each scan's own ceil(ZN / 64), or
  // kZgWindow for the windowed form (scans with columns taller than kWinMaxZ: their own)
  int fpGrid = 0, zg = 0, bpGrid = 0;  // grids: persistent workgroups, <= 0 one per item
  int fpStrided = 0, bpStrided = 0;
  hipStream_t st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool useGraph = true;
  bool fromCache = false;
  double setup_ms = 0, tune_ms = 0, capture_ms = 0;
  ~ctr_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {

constexpr int kZgWindow = 5;

// back-projection class of a scan (0: windowed)
int scan_zg(const ctr_batch& B, const Plan& P) {
  if (B.zg == kZgWindow && P.d.ZN <= kWinMaxZ) return 0;
  return (B.zg > 0 && B.zg < kZgWindow) ? B.zg : std::min(4, (P.d.ZN + 63) / 64);
}

// Descriptor tables for the current tuning (zg and colw change the work lists).
template <typename T>
int build_tables(ctr_batch& B) {
  const int n = B.n, S = B.osNum;
  std::vector<FpBatchItem<T>> fp((size_t)S * n);
  std::vector<ReduceItem> red((size_t)S * n);
  B.fpTotal.assign(S, 0);
  for (int c = 0; c <= 4; ++c) {
    B.clsScans[c].clear();
    B.bpTotal[c].assign(S, 0);
  }
  B.winZ = 0;
  B.winDNV = 0;
  for (int i = 0; i < n; ++i) {
    const int c = scan_zg(B, *B.scans[i]->P);
    B.clsScans[c].push_back(i);
    if (c == 0) {
      B.winZ = std::max(B.winZ, B.scans[i]->P->d.ZN);
      B.winDNV = std::max(B.winDNV, B.scans[i]->P->d.DNV);
    }
  }
  auto up = [&](Mem& m, const void* src, size_t bytes) -> int {
    if (m.p) {
      (void)hipFree(m.p);
      m.p = nullptr;
    }
    int rc = m.alloc(bytes);
    if (rc) return rc;
    CTR_HIP(hipMemcpy(m.p, src, bytes, hipMemcpyHostToDevice));
    return CTR_OK;
  };
  for (int s = 0; s < S; ++s) {
    int fb = 0;
    for (int i = 0; i < n; ++i) {
      Scan& c = *B.scans[i];
      const Plan& P = *c.P;
      const int vc = std::max(vcount(P, s, S), 0);
      const T* vol = B.bf ? (const T*)c.xs.p : (const T*)c.x;
      FpBatchItem<T>& f = fp[(size_t)s * n + i];
      f.a = fp_args<T, T>(P, vol, (T*)c.tp.p, s, S, c.prj,
                          c.rowInv.f() + c.rowOff[s] * c.frame, c.partials.d());
      f.begin = fb;
      f.upv = fp_blocks_per_view(P, B.colw);
      fb += f.upv * vc;
      red[(size_t)s * n + i] = ReduceItem{c.partials.d(), f.upv * vc,
                                          B.sums.d() + ((size_t)i * S + s) * 2};
    }
    B.fpTotal[s] = fb;
  }
  int rc;
  for (int cz = 0; cz <= 4; ++cz) {
    const std::vector<int>& L = B.clsScans[cz];
    if (L.empty()) continue;
    const size_t m = L.size();
    std::vector<BpBatchItem<T>> bp((size_t)S * m), tbp((size_t)S * m);
    for (int s = 0; s < S; ++s) {
      int bb = 0;
      for (size_t j = 0; j < m; ++j) {
        Scan& c = *B.scans[L[j]];
        const Plan& P = *c.P;
        const int vc = std::max(vcount(P, s, S), 0);
        BpBatchItem<T>& b = bp[(size_t)s * m + j];
        b.a = bp_args<T>(P, (const T*)c.tp.p, c.x, s, S, vc, c.colInv.f() + c.NV * s,
                         B.bf ? c.xs.h() : nullptr);
        b.begin = bb;
        b.colBlocks = (P.d.XN * P.d.YN + 3) / 4;
        const int zblocks = cz == 0 ? 1 : (P.d.ZN + 64 * cz - 1) / (64 * cz);
        bb += vc > 0 ? b.colBlocks * zblocks : 0;
        BpBatchItem<T>& t = tbp[(size_t)s * m + j];
        t = b;
        t.a.invCol = B.zeros.f();  // x += acc * 0: timing runs leave the iterate unchanged
      }
      B.bpTotal[cz][s] = bb;
    }
    if ((rc = up(B.bpItems[cz], bp.data(), bp.size() * sizeof(bp[0])))) return rc;
    if ((rc = up(B.tuneBpItems[cz], tbp.data(), tbp.size() * sizeof(tbp[0])))) return rc;
  }
  if ((rc = up(B.fpItems, fp.data(), fp.size() * sizeof(fp[0])))) return rc;
  return up(B.redItems, red.data(), red.size() * sizeof(red[0]));
}

int build(ctr_batch& B) { return B.bf ? build_tables<uint16_t>(B) : build_tables<float>(B); }

template <typename T>
int step_fp(ctr_batch& B, int s) {
  return launch_fp_batch<T>((const FpBatchItem<T>*)B.fpItems.p + (size_t)s * B.n, B.n,
                            B.fpTotal[s], B.colw, B.fpGrid, B.fpStrided, B.st);
}

template <typename T>
int step_bp(ctr_batch& B, int s, bool tuning) {
  for (int cz = 0; cz <= 4; ++cz) {
    const size_t m = B.clsScans[cz].size();
    if (!m) continue;
    const Mem& mem = tuning ? B.tuneBpItems[cz] : B.bpItems[cz];
    const BpBatchItem<T>* items = (const BpBatchItem<T>*)mem.p + (size_t)s * m;
    int rc = cz == 0 ? launch_bp_batch_win<T>(items, (int)m, B.bpTotal[0][s], B.winZ,
                                              B.winDNV, B.bpGrid, B.bpStrided, B.st)
                     : launch_bp_batch<T>(items, (int)m, B.bpTotal[cz][s], cz, B.bpGrid,
                                          B.bpStrided, B.st);
    if (rc) return rc;
  }
  return CTR_OK;
}

int sweep_launches(ctr_batch& B) {
  for (int s = 0; s < B.osNum; ++s) {
    int rc = B.bf ? step_fp<uint16_t>(B, s) : step_fp<float>(B, s);
    if (rc) return rc;
    if ((rc = launch_reduce_batch((const ReduceItem*)B.redItems.p + (size_t)s * B.n, B.n, B.st)))
      return rc;
    if ((rc = B.bf ? step_bp<uint16_t>(B, s, false) : step_bp<float>(B, s, false))) return rc;
  }
  return CTR_OK;
}

// Mean duration (ms) of reps launches of f after one warm-up launch.
template <typename F>
int time_launches(ctr_batch& B, F f, int reps, float* ms) {
  hipEvent_t a, b;
  CTR_HIP(hipEventCreate(&a));
  CTR_HIP(hipEventCreate(&b));
  int rc = f();
  if (!rc) {
    (void)hipEventRecord(a, B.st);
    for (int r = 0; r < reps && !rc; ++r) rc = f();
    (void)hipEventRecord(b, B.st);
    (void)hipEventSynchronize(b);
    float t = 0;
    (void)hipEventElapsedTime(&t, a, b);
    *ms = t / reps;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

void drop_graph(ctr_batch& B) {
  if (B.exec) (void)hipGraphExecDestroy(B.exec);
  if (B.graph) (void)hipGraphDestroy(B.graph);
  B.exec = nullptr;
  B.graph = nullptr;
}

// ---- the tuning cache: one line per batch, "<key>\t<fp_colw fp_grid fp_strided bp_zg
// bp_grid bp_strided>"; the last line of a key wins.  Path: $CTRECON_TUNE_CACHE, else
// $XDG_CACHE_HOME/ctrecon_amd/batch_tuning.txt, else $HOME/.cache/ctrecon_amd/...;
// a cache that cannot be read or written only means tuning again.
std::string cache_path() {
  if (const char* p = std::getenv("CTRECON_TUNE_CACHE")) return p;
  std::string dir;
  if (const char* x = std::getenv("XDG_CACHE_HOME")) dir = x;
  else if (const char* h = std::getenv("HOME")) dir = std::string(h) + "/.cache";
  else return "";
  (void)::mkdir(dir.c_str(), 0755);  // an existing folder is fine
  dir += "/ctrecon_amd";
  (void)::mkdir(dir.c_str(), 0755);
  return dir + "/batch_tuning.txt";
}

// The device (architecture, CU count) and every scan's sizes, voxel sizes, source z
// range and the subset count: what the launch costs depend on.
std::string cache_key(const ctr_batch& B) {
  hipDeviceProp_t prop;
  std::ostringstream k;
  if (hipGetDeviceProperties(&prop, B.device) == hipSuccess)
    k << prop.gcnArchName << "/" << prop.multiProcessorCount;
  k << "|os" << B.osNum << (B.bf ? "|bf16" : "|f32");
  char buf[256];
  for (const auto& c : B.scans) {
    const Plan& P = *c->P;
    std::snprintf(buf, sizeof buf, "|%dx%dx%d,%dx%d,%d,%.5g,%.5g,%.5g,%.5g", P.d.XN, P.d.YN,
                  P.d.ZN, P.d.DNU, P.d.DNV, P.d.PN, P.d.dx, P.d.dz, P.hg.sz.front(),
                  P.hg.sz.back());
    k << buf;
  }
  return k.str();
}

bool cache_load(ctr_batch& B, const std::string& key) {
  const std::string path = cache_path();
  if (path.empty()) return false;
  std::ifstream f(path);
  std::string line, found;
  while (std::getline(f, line)) {
    const size_t tab = line.find('\t');
    if (tab != std::string::npos && line.compare(0, tab, key) == 0 && tab == key.size())
      found = line.substr(tab + 1);
  }
  if (found.empty()) return false;
  int v[6];
  std::istringstream in(found);
  for (int& x : v)
    if (!(in >> x)) return false;
  if (v[3] < 0 || v[3] > kZgWindow || v[1] < 0 || v[4] < 0) return false;
  B.colw = v[0] != 0;
  B.fpGrid = v[1];
  B.fpStrided = v[2] != 0;
  B.zg = v[3];
  B.bpGrid = v[4];
  B.bpStrided = v[5] != 0;
  return true;
}

void cache_store(const ctr_batch& B, const std::string& key) {
  const std::string path = cache_path();
  if (path.empty()) return;
  std::ofstream f(path, std::ios::app);
  if (f)
    f << key << '\t' << (B.colw ? 1 : 0) << ' ' << B.fpGrid << ' ' << B.fpStrided << ' ' << B.zg
      << ' ' << B.bpGrid << ' ' << B.bpStrided << '\n';
}

int autotune(ctr_batch& B) {
  bool all64 = true;
  int maxZN = 0;
  for (auto& c : B.scans) {
    all64 = all64 && c->P->d.DNV == 64;
    maxZN = std::max(maxZN, c->P->d.ZN);
  }
  // schedules: (grid, strided); grid 0 = one workgroup per item
  const int sched[4][2] = {{4 * B.ncu, 0}, {4 * B.ncu, 1}, {8 * B.ncu, 1}, {0, 0}};
  float best = 1e30f;
  bool bcolw = false;
  int bgrid = 0, bstr = 0, rc;
  for (int cw = all64 ? 1 : 0; cw >= 0; --cw) {
    B.colw = cw;
    if ((rc = build(B))) return rc;
    for (const auto& sc : sched) {
      B.fpGrid = sc[0];
      B.fpStrided = sc[1];
      float ms = 0;
      auto f = [&]() { return B.bf ? step_fp<uint16_t>(B, 0) : step_fp<float>(B, 0); };
      if ((rc = time_launches(B, f, 1, &ms))) return rc;
      if (ms < best) {
        best = ms;
        bcolw = cw;
        bgrid = B.fpGrid;
        bstr = B.fpStrided;
      }
    }
  }
  B.colw = bcolw;
  B.fpGrid = bgrid;
  B.fpStrided = bstr;
  best = 1e30f;
  int bzg = 0;
  bgrid = 0;
  bstr = 0;
  const int zgs[5] = {kZgWindow, 0, 1, 2, 4};  // 0: each scan's own ceil(ZN / 64)
  for (int zg : zgs) {
    if (zg > 1 && 64 * (zg / 2) >= maxZN) break;  // more groups than any column needs
    B.zg = zg;
    if ((rc = build(B))) return rc;
    for (const auto& sc : sched) {
      B.bpGrid = sc[0];
      B.bpStrided = sc[1];
      float ms = 0;
      auto f = [&]() {
        return B.bf ? step_bp<uint16_t>(B, 0, true) : step_bp<float>(B, 0, true);
      };
      if ((rc = time_launches(B, f, 1, &ms))) return rc;
      if (ms < best) {
        best = ms;
        bzg = zg;
        bgrid = B.bpGrid;
        bstr = B.bpStrided;
      }
    }
  }
  B.zg = bzg;
  B.bpGrid = bgrid;
  B.bpStrided = bstr;
  return build(B);
}

}  // namespace

extern "C" {

int ctr_batch_create(ctr_plan* const* plans, int nscan, float* const* d_x,
                     const float* const* d_prj, int osNum, int bp_kind, ctr_batch** out) {
  if (!out) return fail(CTR_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (!plans || !d_x || !d_prj || nscan < 1) return fail(CTR_ERR_ARG, "empty batch");
  if (bp_kind != CTR_BP_REFERENCE)
    return fail(CTR_ERR_UNSUPPORTED, "batched OS-SART runs the reference back projector");
  std::unique_ptr<ctr_batch> B(new ctr_batch());
  B->n = nscan;
  B->osNum = osNum;
  B->kind = bp_kind;
  for (int i = 0; i < nscan; ++i) {
    if (!plans[i] || !plans[i]->impl || !d_x[i] || !d_prj[i])
      return fail(CTR_ERR_ARG, "NULL plan or buffer in the batch");
    Plan* P = plans[i]->impl;
    if (P->storage == CTR_STORAGE_BF16_SINO)
      return fail(CTR_ERR_UNSUPPORTED, "batched OS-SART runs with F32 or BF16 storage");
    if (i == 0) {
      B->device = P->device;
      B->bf = P->storage == CTR_STORAGE_BF16;
    } else if (P->device != B->device || (P->storage == CTR_STORAGE_BF16) != B->bf) {
      return fail(CTR_ERR_ARG, "the plans of a batch must share device and storage");
    }
    if (P->method != CTR_METHOD_DIRECT)
      return fail(CTR_ERR_UNSUPPORTED, "batched OS-SART uses the direct projector method");
    if (osNum < 1 || osNum > P->d.PN) return fail(CTR_ERR_ARG, "osNum must be in [1, PN]");
  }
  CTR_HIP(hipSetDevice(B->device));
  (void)hipDeviceGetAttribute(&B->ncu, hipDeviceAttributeMultiprocessorCount, B->device);
  CTR_HIP(hipStreamCreate(&B->st));
  int rc;
  hipEvent_t e0, e1;
  CTR_HIP(hipEventCreate(&e0));
  CTR_HIP(hipEventCreate(&e1));
  (void)hipEventRecord(e0, B->st);
  size_t maxNV = 0;
  for (int i = 0; i < nscan; ++i) {
    auto c = std::make_unique<Scan>();
    Plan& P = *plans[i]->impl;
    c->P = &P;
    c->x = d_x[i];
    c->prj = d_prj[i];
    c->NV = (size_t)P.d.XN * P.d.YN * P.d.ZN;
    c->frame = (size_t)P.d.DNU * P.d.DNV;
    maxNV = std::max(maxNV, c->NV);
    const int maxV = vcount(P, 0, osNum);
    const int upv = std::max(fp_blocks_per_view(P, true), fp_blocks_per_view(P, false));
    if ((rc = c->rowInv.alloc(c->frame * P.d.PN * sizeof(float)))) return rc;
    if ((rc = c->colInv.alloc(c->NV * osNum * sizeof(float)))) return rc;
    if ((rc = c->tp.alloc(c->frame * maxV * (B->bf ? 2 : 4)))) return rc;
    if (B->bf && (rc = c->xs.alloc(c->NV * 2))) return rc;
    if ((rc = c->partials.alloc((size_t)2 * upv * maxV * sizeof(double)))) return rc;
    // row / column sums of the scan's subsets (os_sart.m:14-18), evaluated as the
    // per-scan solver does (without arrays of ones while the columns fit the LDS limit)
    Mem ones;
    const bool ones_bp = P.d.ZN <= kWinMaxZ;
    if (!ones_bp) {
      const size_t no = c->NV;
      if ((rc = ones.alloc(no * sizeof(float)))) return rc;
      if ((rc = launch_fill(ones.f(), no, 1.0f, B->st))) return rc;
    }
    c->rowOff.resize(osNum);
    size_t off = 0;
    for (int s = 0; s < osNum; ++s) {
      const int vc = vcount(P, s, osNum);
      c->rowOff[s] = off;
      float* rI = c->rowInv.f() + off * c->frame;
      if ((rc = launch_fp_ones(P, rI, s, osNum, vc, B->st))) return rc;
      if ((rc = launch_inv_guard(rI, rI, c->frame * vc, B->st))) return rc;
      float* cI = c->colInv.f() + c->NV * s;
      rc = ones_bp ? launch_bp_ones(P, cI, s, osNum, vc, B->st)
                   : launch_bp<float>(P, ones.f(), cI, s, osNum, vc, CTR_BP_REFERENCE, B->st,
                                      nullptr, 0, nullptr);
      if (rc) return rc;
      if ((rc = launch_inv_guard(cI, cI, c->NV, B->st))) return rc;
      off += vc;
    }
    CTR_HIP(hipStreamSynchronize(B->st));  // before `ones` is freed
    B->scans.push_back(std::move(c));
  }
  (void)hipEventRecord(e1, B->st);
  CTR_HIP(hipEventSynchronize(e1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  B->setup_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if ((rc = B->sums.alloc((size_t)nscan * osNum * 2 * sizeof(double)))) return rc;
  if ((rc = B->zeros.alloc(maxNV * sizeof(float)))) return rc;
  if ((rc = launch_fill(B->zeros.f(), maxNV, 0.0f, B->st))) return rc;
  if (B->bf)
    for (auto& c : B->scans)
      if ((rc = launch_to_bf16_shadow(c->xs.h(), c->x, c->NV, B->st))) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  const std::string key = cache_key(*B);
  if (cache_load(*B, key)) {
    B->fromCache = true;
    if ((rc = build(*B))) return rc;
  } else {
    if ((rc = autotune(*B))) return rc;
    cache_store(*B, key);
  }
  CTR_HIP(hipStreamSynchronize(B->st));
  B->tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  *out = B.release();
  return CTR_OK;
}

int ctr_batch_set_graph(ctr_batch* B, int enable) {
  if (!B) return fail(CTR_ERR_ARG, "NULL batch");
  B->useGraph = enable != 0;
  return CTR_OK;
}

int ctr_batch_set_tuning(ctr_batch* B, int fp_colw, int fp_grid, int fp_strided, int bp_zg,
                         int bp_grid, int bp_strided) {
  if (!B) return fail(CTR_ERR_ARG, "NULL batch");
  if (bp_zg < 0 || bp_zg > kZgWindow || fp_grid < 0 || bp_grid < 0)
    return fail(CTR_ERR_ARG, "bad tuning");
  if (fp_colw)
    for (auto& c : B->scans)
      if (c->P->d.DNV != 64) return fail(CTR_ERR_ARG, "column-per-wave mapping needs DNV == 64");
  CTR_HIP(hipSetDevice(B->device));
  B->colw = fp_colw != 0;
  B->fpGrid = fp_grid;
  B->fpStrided = fp_strided != 0;
  B->zg = bp_zg;
  B->bpGrid = bp_grid;
  B->bpStrided = bp_strided != 0;
  drop_graph(*B);
  return build(*B);
}

int ctr_batch_get_info(const ctr_batch* B, ctr_batch_info* info) {
  if (!B || !info) return fail(CTR_ERR_ARG, "NULL argument");
  info->fp_colw = B->colw;
  info->fp_grid = B->fpGrid;
  info->fp_strided = B->fpStrided;
  info->bp_zg = B->zg;
  info->bp_grid = B->bpGrid;
  info->bp_strided = B->bpStrided;
  info->setup_ms = B->setup_ms;
  info->tune_ms = B->tune_ms;
  info->capture_ms = B->capture_ms;
  info->from_cache = B->fromCache ? 1 : 0;
  long long f = 0, b = 0;
  for (int s = 0; s < B->osNum; ++s) {
    f += B->fpTotal[s];
    for (int cz = 0; cz <= 4; ++cz)
      if (!B->bpTotal[cz].empty()) b += B->bpTotal[cz][s];
  }
  info->fp_items = f;
  info->bp_items = b;
  return CTR_OK;
}

int ctr_batch_sweeps(ctr_batch* B, int maxIter, double stop_resid, double* resid,
                     double* sweep_ms, int* done) {
  if (!B) return fail(CTR_ERR_ARG, "NULL batch");
  if (maxIter < 0) return fail(CTR_ERR_ARG, "maxIter must be >= 0");
  CTR_HIP(hipSetDevice(B->device));
  int rc;
  // the caller may have changed the iterates: refresh the BF16 shadows
  if (B->bf)
    for (auto& c : B->scans)
      if ((rc = launch_to_bf16_shadow(c->xs.h(), c->x, c->NV, B->st))) return rc;
  if (B->useGraph && !B->exec) {
    const auto t0 = std::chrono::steady_clock::now();
    CTR_HIP(hipStreamBeginCapture(B->st, hipStreamCaptureModeThreadLocal));
    rc = sweep_launches(*B);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(B->st, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(CTR_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    B->graph = g;
    CTR_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
    B->capture_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  hipEvent_t e0, e1;
  CTR_HIP(hipEventCreate(&e0));
  CTR_HIP(hipEventCreate(&e1));
  std::vector<double> h((size_t)B->n * B->osNum * 2);
  int it = 0;
  rc = CTR_OK;
  for (; it < maxIter; ++it) {
    (void)hipEventRecord(e0, B->st);
    if (B->useGraph) {
      if (hipGraphLaunch(B->exec, B->st) != hipSuccess) {
        rc = fail(CTR_ERR_HIP, "hipGraphLaunch failed");
        break;
      }
    } else if ((rc = sweep_launches(*B))) {
      break;
    }
    (void)hipEventRecord(e1, B->st);
    if (hipMemcpyAsync(h.data(), B->sums.p, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                       B->st) != hipSuccess ||
        hipStreamSynchronize(B->st) != hipSuccess) {
      rc = fail(CTR_ERR_HIP, "residual readback failed");
      break;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (sweep_ms) sweep_ms[it] = ms;
    bool all = true;
    for (int i = 0; i < B->n; ++i) {
      double rr = 0, bb = 0;
      for (int s = 0; s < B->osNum; ++s) {
        rr += h[((size_t)i * B->osNum + s) * 2];
        bb += h[((size_t)i * B->osNum + s) * 2 + 1];
      }
      const double r = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
      if (resid) resid[(size_t)it * B->n + i] = r;
      all = all && r <= stop_resid;
    }
    if (stop_resid > 0 && all) {
      ++it;
      break;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (done) *done = it;
  return rc;
}

int ctr_batch_destroy(ctr_batch* B) {
  if (B) {
    (void)hipSetDevice(B->device);
    (void)hipStreamSynchronize(B->st);
    delete B;
  }
  return CTR_OK;
}

}  // extern "C"
