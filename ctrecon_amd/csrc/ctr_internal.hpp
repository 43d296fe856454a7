// Internal declarations shared by the host plan code and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ctrecon.h"

namespace ctr {

// Largest column height (voxels) the windowed back projector keeps in LDS:
// 4 waves x kWinMaxZ floats = 32 KiB per workgroup.
constexpr int kWinMaxZ = 2048;

// Per-view constants (computed on the host in FP64, stored FP32).
//   (sx, sy, sz)  rotated source position (DD3_GPU.cu:875-878, 1441-1444)
//   az            (z0 + zshift)/dz + ocz: z index of the source plane offset
//   xmarch        1 when |cos| <= |sin|: march x-slices (DD3_GPU.cu:914)
//   bdir          +1 / -1: direction in which the in-plane boundary positions of
//                 this view increase with the boundary index
//   wx            1 when |src.x| > |src.y| (FP64 decision): the reference back
//                 projector then uses the x component as the cell width direction
//                 (DD3_GPU.cu:1453-1460)
//   azi, azf      az split into the nearest integer and the rest (|azf| <= 1/2): the
//                 forward projector and its transpose place z footprints relative to
//                 azi, where FP32 keeps 16x or more finer positions than at az
struct ViewP {
  float sx, sy, sz, az;
  int xmarch, bdir, wx, azi;
  float azf;
  int pad0, pad1, pad2;
};

// Constants of the reference back projector (DD3_GPU.cu:1428-1437, calDetParas
// :1189-1217), all in unpadded cell-centre index units.
struct BpConsts {
  float S;        // S2D / ddv
  float ib;       // 1 / dbeta
  float hz;       // dz / 2
  float inv_sid;  // 1 / |(x0, y0)|
  float cU, cV;   // detector centre indices (detCtrIdxU/V)
  float dx, dz;
};

// FP64 copies of what the reference back projector needs for the in-plane
// (fan-angle) coordinate, which must be resolved to well below a detector cell
// on detectors several hundred cells wide.
struct BpConsts64 {
  double ib, cU, inv_sid, dx, ocx, ocy;
  double S, cV, ocz;  // S2D / ddv, detCtrIdxV, objCntIdx.z (the z-line back projector)
};

// Geometry constants every kernel needs.
struct Dims {
  int XN, YN, ZN, DNU, DNV, PN;
  float ocx, ocy, ocz;  // objCntIdx (DD3_GPU.cu:1029-1031)
  float dx, dz;
  float beta0, dbeta_g;  // fan angle of boundary 0 and mean boundary spacing (adjoint guess)
  float B0, dB;          // row boundary index of row 0 and mean spacing (adjoint guess)
  float dBmin;           // smallest row spacing of the FP32 boundary table
};

// Host copies used by host-side accounting (cone voxel counts).
struct HostGeo {
  std::vector<double> sz;     // source z per view
  std::vector<uint8_t> mask;  // 1 where the mask byte is non-zero
  double ocx, ocy, ocz, dx, dz, S2D, bz0, bzN;
};

// Operands of one forward-projection launch over one scan (TV / TO: storage types of vol
// and out, float or uint16_t holding BF16).  b / invRow / partials are set for the fused
// OS-SART residual and NULL otherwise.
template <typename TV, typename TO = TV>
struct FpArgs {
  const TV* vol;
  TO* out;
  const ViewP* views;
  const float2* bnd;
  const double2* bnd64;
  const float4* cols;
  const float* Bz;
  const float* h2;
  const float* invdB;
  const float* maskf;
  const float2* mext;
  Dims d;
  int vb, vs;
  const float* b;
  const float* invRow;
  double* partials;
};

// Operands of one reference back-projection launch over one scan (T: storage type of
// prj; vol FP32).  invCol / shadow are set for the OS-SART update and NULL otherwise.
template <typename T>
struct BpArgs {
  const T* prj;
  float* vol;
  const ViewP* views;
  const double2* src64;
  const uint8_t* mask;
  Dims d;
  BpConsts c;
  BpConsts64 c64;
  int vb, vs, vc;
  const float* invCol;
  uint16_t* shadow;
};

// Batched launches: one work list over several scans.  Work items [begin, begin + n)
// of the flat list belong to the scan; upv = blocks per view (forward) and
// colBlocks = blocks of 4 voxel columns per z block (back).
template <typename T>
struct FpBatchItem {
  FpArgs<T> a;
  int begin, upv;
};
template <typename T>
struct BpBatchItem {
  BpArgs<T> a;
  int begin, colBlocks;
};
// Per-scan reduction of the fused residual partials of one batched subset step.
struct ReduceItem {
  const double* partials;
  int n;
  double* out;
};

struct Plan {
  int device = 0;
  Dims d{};
  BpConsts bc{};
  BpConsts64 bc64{};
  // device tables
  ViewP* views = nullptr;      // [PN]
  double2* src64 = nullptr;    // [PN]         rotated source (x, y) in FP64
  float2* bnd = nullptr;       // [PN][DNU+1]  (a_b, f_b): t_b(ii) = a_b + ii*f_b
  double2* bnd64 = nullptr;    // [PN][DNU+1]  the same in FP64 (forward projector, transpose)
  float4* cols = nullptr;      // [PN][DNU]    (r0, r1, G, Q)
  float* Bz = nullptr;         // [DNV+1]      (bz_v - z0)/dz
  float* h2 = nullptr;         // [DNV]        (zds_v - z0)^2
  float* invdB = nullptr;      // [DNV]        1/(Bz[v+1]-Bz[v])
  double2* rot64 = nullptr;    // [PN]         (cos, sin) of the regularised view angle, FP64
  double* sz64 = nullptr;      // [PN]         source z (z0 + zshift), FP64
  double2* det64 = nullptr;    // [DNU]        detector cell centres (xds, yds), FP64
  float* zc = nullptr;         // [DNV]        (zds_v - z0)/dz
  uint8_t* mask = nullptr;     // [XN*YN]
  float* maskf = nullptr;      // [XN*YN]      mask byte as float (the forward multiplier)
  float2* mext = nullptr;      // [XN + YN]    clip interval of the mask per x- then y-slice
  // host copies (for culling and solvers)
  std::vector<ViewP> h_views;
  HostGeo hg;
  // projector method: CTR_METHOD_DIRECT, CTR_METHOD_SAT or CTR_METHOD_PSEUDO_DD
  int method = CTR_METHOD_DIRECT;
  // storage of the solvers' projector operands: CTR_STORAGE_F32, CTR_STORAGE_BF16 or
  // CTR_STORAGE_BF16_SINO
  int storage = CTR_STORAGE_F32;
  // scratch
  float* sat_vol = nullptr;         // volume integral images (hi then lo planes)
  size_t sat_vol_n = 0;
  float* sat_sino = nullptr;        // sinogram integral images of one launch
  size_t sat_sino_n = 0;
  float* scratch_prj = nullptr;     // [max_views][DNU][DNV]
  size_t scratch_prj_views = 0;
  double* partials = nullptr;       // reduction partials
  double* zl_cpre = nullptr;        // z-line back projector: row prefix of one launch's views
  size_t zl_cpre_n = 0;
  size_t partials_n = 0;
  // the plan's own stream (default priority, synchronises with the null stream): the
  // OS-SART / CG drivers, the residual and the host-buffer entry points run on it
  hipStream_t stream = nullptr;
  // conjugate-gradient vectors kept between calls: r, p, t, the sinogram between the
  // projectors, the BF16 copy of the operator's input; and the residual's two sums
  void* cg[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  size_t cg_bytes[5] = {0, 0, 0, 0, 0};
  double* sums2 = nullptr;
};

// Every device allocation of the library goes through dev_malloc (hipMalloc, with the
// allocation-failure injection of ctr_debug_fail_alloc_after).
hipError_t dev_malloc(void** p, size_t bytes);
// *p holds at least `need` bytes afterwards (reallocated, contents lost, when smaller).
int ensure_bytes(void** p, size_t* have, size_t need, const char* what);

// Relative data residual ||A x - b|| / ||b|| over all views (FP64 sums, added in view
// order), forward-projecting `chunk` views at a time into prj ([chunk][DNU][DNV]) with
// partials (2 fp_partial_blocks(P, chunk) doubles) and sums (kResidualSums doubles) as
// scratch, on stream st; synchronous.  Allocates nothing for the direct method.
constexpr int kResidualSums = 128;
int residual_run(Plan& P, const float* d_x, const float* d_prj, float* prj, int chunk,
                 double* partials, double* sums, hipStream_t st, double* out);

// error handling
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define CTR_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::ctr::fail(CTR_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// kernel launchers (kernels.hip).  T is the storage type of the projector's input and
// output arrays: float, or uint16_t holding BF16 (FP32 accumulation either way).
// shadow: when not NULL, an UPDATE launch also writes the updated volume as BF16.
template <typename TV, typename TO>
int launch_fp(const Plan& P, const TV* vol, TO* out, int vb, int vs, int vc, hipStream_t st,
              const float* b, const float* invRow, double* partials);
template <typename T>
int launch_bp(const Plan& P, const T* prj, float* vol, int vb, int vs, int vc, int kind,
              hipStream_t st, const float* invCol, int accumulate, uint16_t* shadow);
// A 1 and B 1 (the OS-SART row and column sums) of the branchless model over the views
// vb + k vs, k < vc, without reading a volume or sinogram of ones; launch_bp_ones needs
// ZN <= kWinMaxZ.
int launch_fp_ones(const Plan& P, float* out, int vb, int vs, int vc, hipStream_t st);
int launch_bp_ones(const Plan& P, float* vol, int vb, int vs, int vc, hipStream_t st);
template <typename T>
int launch_scale_matched(const Plan& P, const T* in, float* out, int vb, int vs, int vc,
                         hipStream_t st);
template <typename TV, typename TO>
FpArgs<TV, TO> fp_args(const Plan& P, const TV* vol, TO* out, int vb, int vs, const float* b,
                       const float* invRow, double* partials);
template <typename T>
BpArgs<T> bp_args(const Plan& P, const T* prj, float* vol, int vb, int vs, int vc,
                  const float* invCol, uint16_t* shadow);
int fp_blocks_per_view(const Plan& P, bool colw);
// batched launches: grid = number of persistent workgroups (<= 0: one per work item),
// strided = interleaved instead of contiguous item ranges
template <typename T>
int launch_fp_batch(const FpBatchItem<T>* items, int nscan, int total, bool colw, int grid,
                    int strided, hipStream_t st);
template <typename T>
int launch_bp_batch(const BpBatchItem<T>* items, int nscan, int total, int zg, int grid,
                    int strided, hipStream_t st);
// windowed form: one work item per block of 4 columns over all z; zmax = tallest column,
// maxDNV = most detector rows of a scan in the list
template <typename T>
int launch_bp_batch_win(const BpBatchItem<T>* items, int nscan, int total, int zmax,
                        int maxDNV, int grid, int strided, hipStream_t st);
int launch_reduce_batch(const ReduceItem* items, int nscan, hipStream_t st);
int launch_to_bf16(uint16_t* dst, const float* src, size_t n, hipStream_t st);
// the solvers' dithered BF16 shadow of an iterate (to_bf16_dither)
int launch_to_bf16_shadow(uint16_t* dst, const float* src, size_t n, hipStream_t st);
int launch_from_bf16(float* dst, const uint16_t* src, size_t n, hipStream_t st);
int launch_mask_volume(const Plan& P, float* vol, hipStream_t st);
int launch_fill(float* p, size_t n, float v, hipStream_t st);
int launch_copy(float* dst, const float* src, size_t n, hipStream_t st);
int launch_inv_guard(const float* in, float* out, size_t n, hipStream_t st);
int launch_axpy(float* y, const float* x, double a, size_t n, hipStream_t st);
int launch_xpby(float* p, const float* r, double b, size_t n, hipStream_t st);
int launch_sub(float* out, const float* a, const float* b, size_t n, hipStream_t st);
int launch_dot(const float* a, const float* b, size_t n, double* partials, int nparts,
               double* result, hipStream_t st);
int launch_reduce_pairs(const double* partials, int nparts, double* result, hipStream_t st);
int fp_partial_blocks(const Plan& P, int vc);
// pseudo distance-driven pair (pdd.hip), FP32 operands
int launch_fp_pdd(const Plan& P, const float* vol, float* out, int vb, int vs, int vc,
                  hipStream_t st);
int launch_bp_pdd(const Plan& P, const float* prj, float* vol, int vb, int vs, int vc,
                  hipStream_t st);
// z-line back projector (pdd.hip); cpre: zline_prefix_doubles(d, vc) doubles of scratch
size_t zline_prefix_doubles(const Dims& d, int vc);
int launch_bp_zline(const Plan& P, const float* prj, double* cpre, float* vol, int vb, int vs,
                    int vc, hipStream_t st);
// integral-image method (sat.hip)
size_t sat_volume_floats(const Dims& d);
size_t sat_sino_floats(const Dims& d, int vc);
template <typename T>
int launch_volume_sat(const Plan& P, const T* vol, float* sat, hipStream_t st);
template <typename TO>
int launch_fp_sat(const Plan& P, const float* sat, TO* out, int vb, int vs, int vc,
                  hipStream_t st, const float* b, const float* invRow, double* partials);
template <typename T>
int launch_bp_sat(const Plan& P, const T* prj, float* sat, float* vol, int vb, int vs, int vc,
                  hipStream_t st, const float* invCol, int accumulate, uint16_t* shadow);

}  // namespace ctr

// The opaque handle of the C ABI.
struct ctr_plan {
  ctr::Plan* impl;
};
