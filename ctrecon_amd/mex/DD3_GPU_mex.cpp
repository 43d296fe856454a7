// DD3_GPU MEX gateway over the C ABI of libctrecon_amd.so.
//
// Drop-in for the reference gateway (iterativeRecon/DD3_GPU.cu:1864-2066): the same
// name, actions and right-hand sides, so iterativeRecon/dd3.m runs unchanged:
//
//   sino = DD3_GPU('Proj', x0, y0, z0, DNU, DNV, xds, yds, zds, xoff, yoff, zoff,
//                  angles, zshifts, PN, XN, YN, ZN, img, dx, dz, mask, gpu, prjMode)   % 24 rhs
//   img  = DD3_GPU('Back', ... same ..., sino, dx, dz, mask, gpu, bp_type, prjMode)  % 25 rhs
//
// and two actions that run a whole loop on the GPU (no reference counterpart; the
// reference loops in MATLAB, iterativeRecon/os_sart.m:21-44, conjgrad.m:8-34):
//
//   [x, resid] = DD3_GPU('OSSART', <the 17 geometry rhs of 'Proj'>, x0img, prj, dx, dz,
//                        mask, gpu, osNum, iterNum, bpKind)
//   [x, resid] = DD3_GPU('CG',     <the 17 geometry rhs of 'Proj'>, x0img, prj, dx, dz,
//                        mask, gpu, iterNum, bpKind)
//
// Differences from the reference: arguments are checked (class and size) and errors
// are raised with mexErrMsgIdAndTxt instead of printed; the input image is not
// masked in place; the 'Back' output has the dimensions the kernel writes,
// {ZN, XN, YN} (the reference declares {ZN, YN, XN}, DD3_GPU.cu:1969, which agrees
// only for square slices).
#include <cstdint>
#include <cstring>
#include <string>

#include "ctrecon.h"
#include "mex.h"

namespace {

[[noreturn]] void fail(const char* id, const std::string& msg) {
  mexErrMsgIdAndTxt(id, "%s", msg.c_str());
  throw 0;  // not reached: mexErrMsgIdAndTxt does not return
}

float get_f(const mxArray* a, const char* name) {
  if (!a || !mxIsSingle(a) || mxGetNumberOfElements(a) < 1)
    fail("DD3_GPU:type", std::string(name) + " must be a single scalar");
  return *(const float*)mxGetData(a);
}

int get_i(const mxArray* a, const char* name) {
  if (!a || mxGetNumberOfElements(a) < 1) fail("DD3_GPU:type", std::string(name) + " is empty");
  if (mxIsInt32(a)) return *(const int32_t*)mxGetData(a);
  if (mxIsDouble(a)) return (int)mxGetScalar(a);  // dd3.m passes gpu_id as a double
  fail("DD3_GPU:type", std::string(name) + " must be int32 (or double)");
}

const float* get_fv(const mxArray* a, size_t n, const char* name) {
  if (!a || !mxIsSingle(a) || mxGetNumberOfElements(a) < n)
    fail("DD3_GPU:type", std::string(name) + " must be single with at least " +
                             std::to_string(n) + " elements");
  return (const float*)mxGetData(a);
}

const unsigned char* get_mask(const mxArray* a, size_t n) {
  if (!a || (!mxIsInt8(a) && !mxIsUint8(a)) || mxGetNumberOfElements(a) < n)
    fail("DD3_GPU:type", "mask must be int8 with at least XN*YN elements");
  return (const unsigned char*)mxGetData(a);
}

struct Geo {
  float x0, y0, z0, xoff, yoff, zoff;
  int DNU, DNV, PN, XN, YN, ZN;
  const float *xds, *yds, *zds, *ang, *zsh;
};

// rhs[1..17]: x0 y0 z0 DNU DNV xds yds zds xoff yoff zoff angles zshifts PN XN YN ZN
Geo read_geo(const mxArray* const* a) {
  Geo g;
  g.x0 = get_f(a[1], "x0");
  g.y0 = get_f(a[2], "y0");
  g.z0 = get_f(a[3], "z0");
  g.DNU = get_i(a[4], "DNU");
  g.DNV = get_i(a[5], "DNV");
  if (g.DNU < 1 || g.DNV < 1) fail("DD3_GPU:size", "DNU and DNV must be positive");
  g.xds = get_fv(a[6], g.DNU, "xds");
  g.yds = get_fv(a[7], g.DNU, "yds");
  g.zds = get_fv(a[8], g.DNV, "zds");
  g.xoff = get_f(a[9], "imgXoffset");
  g.yoff = get_f(a[10], "imgYoffset");
  g.zoff = get_f(a[11], "imgZoffset");
  g.PN = get_i(a[14], "PN");
  if (g.PN < 1) fail("DD3_GPU:size", "PN must be positive");
  g.ang = get_fv(a[12], g.PN, "viewangles");
  g.zsh = get_fv(a[13], g.PN, "zshifts");
  g.XN = get_i(a[15], "XN");
  g.YN = get_i(a[16], "YN");
  g.ZN = get_i(a[17], "ZN");
  if (g.XN < 1 || g.YN < 1 || g.ZN < 1) fail("DD3_GPU:size", "XN, YN, ZN must be positive");
  return g;
}

void check(int rc) {
  if (rc != CTR_OK) fail("DD3_GPU:ctr", ctr_last_error());
}

size_t nvox(const Geo& g) { return (size_t)g.XN * g.YN * g.ZN; }
size_t nsino(const Geo& g) { return (size_t)g.DNU * g.DNV * g.PN; }

void proj(mxArray* plhs[], const mxArray* const* a) {
  const Geo g = read_geo(a);
  const float* img = get_fv(a[18], nvox(g), "image");
  const float dx = get_f(a[19], "dx"), dz = get_f(a[20], "dz");
  const unsigned char* mask = get_mask(a[21], (size_t)g.XN * g.YN);
  const int gpu = get_i(a[22], "gpu"), mode = get_i(a[23], "prjMode");
  const mwSize dims[3] = {(mwSize)g.DNV, (mwSize)g.DNU, (mwSize)g.PN};
  plhs[0] = mxCreateNumericArray(3, dims, mxSINGLE_CLASS, mxREAL);
  check(ctr_dd3_proj(g.x0, g.y0, g.z0, g.DNU, g.DNV, g.xds, g.yds, g.zds, g.xoff, g.yoff,
                     g.zoff, g.ang, g.zsh, g.PN, g.XN, g.YN, g.ZN, img,
                     (float*)mxGetData(plhs[0]), dx, dz, mask, gpu, mode));
}

void back(mxArray* plhs[], const mxArray* const* a) {
  const Geo g = read_geo(a);
  const float* sino = get_fv(a[18], nsino(g), "sinogram");
  const float dx = get_f(a[19], "dx"), dz = get_f(a[20], "dz");
  const unsigned char* mask = get_mask(a[21], (size_t)g.XN * g.YN);
  const int gpu = get_i(a[22], "gpu"), squared = get_i(a[23], "bp_type");
  const int mode = get_i(a[24], "prjMode");
  const mwSize dims[3] = {(mwSize)g.ZN, (mwSize)g.XN, (mwSize)g.YN};
  plhs[0] = mxCreateNumericArray(3, dims, mxSINGLE_CLASS, mxREAL);
  check(ctr_dd3_back(g.x0, g.y0, g.z0, g.DNU, g.DNV, g.xds, g.yds, g.zds, g.xoff, g.yoff,
                     g.zoff, g.ang, g.zsh, g.PN, g.XN, g.YN, g.ZN, (float*)mxGetData(plhs[0]),
                     sino, dx, dz, mask, gpu, squared, mode));
}

// Device resources of a loop action.  MATLAB's mexErrMsgIdAndTxt leaves the MEX function
// without running C++ destructors, so they are released explicitly before any error is
// raised; the output arrays are created first (MATLAB frees those itself on an error).
struct LoopRes {
  ctr_plan* plan = nullptr;
  void* x = nullptr;
  void* prj = nullptr;
  void release() {
    if (x) ctr_free(x);
    if (prj) ctr_free(prj);
    if (plan) ctr_plan_destroy(plan);
    x = nullptr;
    prj = nullptr;
    plan = nullptr;
  }
};

void loop(int nlhs, mxArray* plhs[], int nrhs, const mxArray* const* a, bool sart) {
  if (nrhs != (sart ? 27 : 26))
    fail("DD3_GPU:usage", sart ? "OSSART takes 27 arguments" : "CG takes 26 arguments");
  const Geo g = read_geo(a);
  const float* x0 = get_fv(a[18], nvox(g), "initial image");
  const float* prj = get_fv(a[19], nsino(g), "sinogram");
  const float dx = get_f(a[20], "dx"), dz = get_f(a[21], "dz");
  const unsigned char* mask = get_mask(a[22], (size_t)g.XN * g.YN);
  const int gpu = get_i(a[23], "gpu");
  const int osNum = sart ? get_i(a[24], "osNum") : 0;
  const int iterNum = get_i(a[sart ? 25 : 24], "iterNum");
  const int kind = get_i(a[sart ? 26 : 25], "bpKind");
  if (iterNum < 0) fail("DD3_GPU:size", "iterNum must be >= 0");
  ctr_geometry geo{g.x0, g.y0, g.z0, g.DNU, g.DNV, g.xds, g.yds, g.zds, g.xoff, g.yoff, g.zoff,
                   g.ang, g.zsh, g.PN, g.XN, g.YN, g.ZN, dx, dz, mask};
  const mwSize dims[3] = {(mwSize)g.ZN, (mwSize)g.XN, (mwSize)g.YN};
  plhs[0] = mxCreateNumericArray(3, dims, mxSINGLE_CLASS, mxREAL);
  // the residuals: the second output, or a scratch array when it is not requested
  mxArray* res = mxCreateDoubleMatrix(1, iterNum, mxREAL);
  if (nlhs > 1) plhs[1] = res;
  LoopRes R;
  auto ok = [&](int rc) {
    if (rc != CTR_OK) {
      const std::string msg = ctr_last_error();
      R.release();
      fail("DD3_GPU:ctr", msg);
    }
  };
  ok(ctr_plan_create(&geo, gpu, &R.plan));
  ok(ctr_malloc(&R.x, nvox(g) * sizeof(float), gpu));
  ok(ctr_malloc(&R.prj, nsino(g) * sizeof(float), gpu));
  ok(ctr_memcpy_h2d(R.x, x0, nvox(g) * sizeof(float)));
  ok(ctr_memcpy_h2d(R.prj, prj, nsino(g) * sizeof(float)));
  if (sart)
    ok(ctr_os_sart(R.plan, (float*)R.x, (const float*)R.prj, osNum, iterNum, kind, mxGetPr(res)));
  else
    ok(ctr_conjgrad(R.plan, (float*)R.x, (const float*)R.prj, iterNum, kind, mxGetPr(res)));
  ok(ctr_memcpy_d2h(mxGetData(plhs[0]), R.x, nvox(g) * sizeof(float)));
  R.release();
  if (nlhs <= 1) mxDestroyArray(res);
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs < 1 || !mxIsChar(prhs[0]))
    fail("DD3_GPU:usage", "first argument must be 'Proj', 'Back', 'OSSART' or 'CG'");
  char op[16] = {0};
  if (mxGetString(prhs[0], op, sizeof(op))) fail("DD3_GPU:usage", "action name too long");
  if (!std::strcmp(op, "Proj")) {
    if (nrhs != 24) fail("DD3_GPU:usage", "Proj takes 24 arguments (DD3_GPU.cu:2032)");
    proj(plhs, prhs);
  } else if (!std::strcmp(op, "Back")) {
    if (nrhs != 25) fail("DD3_GPU:usage", "Back takes 25 arguments (DD3_GPU.cu:2042)");
    back(plhs, prhs);
  } else if (!std::strcmp(op, "OSSART")) {
    loop(nlhs, plhs, nrhs, prhs, true);
  } else if (!std::strcmp(op, "CG")) {
    loop(nlhs, plhs, nrhs, prhs, false);
  } else {
    fail("DD3_GPU:usage", std::string("unsupported action ") + op);
  }
}
