// Drives the DD3_GPU gateway through the stand-in MATLAB API.  TEST INFRASTRUCTURE.
//   mex_driver errors        -> one line per bad call: the error identifier raised
//   mex_driver run <dir>     -> reads geometry/inputs written by the test, calls
//                               'Proj', 'Back', 'OSSART' and 'CG', writes the outputs
//   mex_driver leak <dir>    -> 'OSSART' with the k-th device allocation failing, for
//                               k = 0, 1, ... until the call succeeds, leaving the gateway
//                               the way MATLAB does (longjmp, no destructors); prints the
//                               plans and device buffers still alive after each call
#include <dlfcn.h>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <set>
#include <string>
#include <vector>

#include "ctrecon.h"
#include "mex.h"

// Live plans and device buffers: the gateway's calls of these C-ABI functions land here
// (the executable's definitions come first) and are forwarded to the library.  Only
// handles created through here are counted (the library's own internal calls may or may
// not pass through).
static std::set<void*> g_plans, g_bufs;
template <typename F>
static F next_fn(const char* name) {
  return reinterpret_cast<F>(dlsym(RTLD_NEXT, name));
}
extern "C" int ctr_plan_create(const ctr_geometry* g, int device, ctr_plan** out) {
  static auto f = next_fn<int (*)(const ctr_geometry*, int, ctr_plan**)>("ctr_plan_create");
  const int rc = f(g, device, out);
  if (rc == CTR_OK && out && *out) g_plans.insert(*out);
  return rc;
}
extern "C" int ctr_plan_destroy(ctr_plan* p) {
  static auto f = next_fn<int (*)(ctr_plan*)>("ctr_plan_destroy");
  g_plans.erase(p);
  return f(p);
}
extern "C" int ctr_malloc(void** p, unsigned long long bytes, int device) {
  static auto f = next_fn<int (*)(void**, unsigned long long, int)>("ctr_malloc");
  const int rc = f(p, bytes, device);
  if (rc == CTR_OK && p && *p) g_bufs.insert(*p);
  return rc;
}
extern "C" int ctr_free(void* p) {
  static auto f = next_fn<int (*)(void*)>("ctr_free");
  g_bufs.erase(p);
  return f(p);
}

static mxArray* f32(float v) {
  const mwSize d[2] = {1, 1};
  mxArray* a = mxCreateNumericArray(2, d, mxSINGLE_CLASS, mxREAL);
  *(float*)mxGetData(a) = v;
  return a;
}
static mxArray* i32(int v) {
  const mwSize d[2] = {1, 1};
  mxArray* a = mxCreateNumericArray(2, d, mxINT32_CLASS, mxREAL);
  *(int*)mxGetData(a) = v;
  return a;
}
static mxArray* f64(double v) {
  mxArray* a = mxCreateDoubleMatrix(1, 1, mxREAL);
  *mxGetPr(a) = v;
  return a;
}
static mxArray* vecf(const std::vector<float>& v) {
  const mwSize d[2] = {1, v.size()};
  mxArray* a = mxCreateNumericArray(2, d, mxSINGLE_CLASS, mxREAL);
  std::memcpy(mxGetData(a), v.data(), v.size() * 4);
  return a;
}
static mxArray* veci8(const std::vector<unsigned char>& v) {
  const mwSize d[2] = {1, v.size()};
  mxArray* a = mxCreateNumericArray(2, d, mxINT8_CLASS, mxREAL);
  std::memcpy(mxGetData(a), v.data(), v.size());
  return a;
}

template <typename T>
static std::vector<T> readv(const std::string& path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  const size_t n = (size_t)f.tellg() / sizeof(T);
  f.seekg(0);
  std::vector<T> v(n);
  f.read((char*)v.data(), n * sizeof(T));
  return v;
}
static void writev(const std::string& path, const void* p, size_t bytes) {
  std::ofstream f(path, std::ios::binary);
  f.write((const char*)p, bytes);
}

struct Call {
  std::vector<mxArray*> rhs;
  ~Call() {
    for (auto a : rhs) mxDestroyArray(a);
  }
};

// rhs[0..17] for a geometry read from dir (or a 2x2x2 dummy)
static void geometry(Call& c, const char* action, const std::string& dir) {
  std::vector<int> n = dir.empty() ? std::vector<int>{2, 2, 2, 2, 2, 2}
                                   : readv<int>(dir + "/sizes.i32");  // DNU DNV PN XN YN ZN
  std::vector<float> sc = dir.empty() ? std::vector<float>{0, 100, 0, 0, 0, 0, 1, 1}
                                      : readv<float>(dir + "/scalars.f32");
  auto arr = [&](const char* name, int len) {
    return dir.empty() ? std::vector<float>(len, 0.f) : readv<float>(dir + "/" + name);
  };
  c.rhs.push_back(mxCreateString(action));
  c.rhs.push_back(f32(sc[0]));
  c.rhs.push_back(f32(sc[1]));
  c.rhs.push_back(f32(sc[2]));
  c.rhs.push_back(i32(n[0]));
  c.rhs.push_back(i32(n[1]));
  c.rhs.push_back(vecf(arr("xds.f32", n[0])));
  c.rhs.push_back(vecf(arr("yds.f32", n[0])));
  c.rhs.push_back(vecf(arr("zds.f32", n[1])));
  c.rhs.push_back(f32(sc[3]));
  c.rhs.push_back(f32(sc[4]));
  c.rhs.push_back(f32(sc[5]));
  c.rhs.push_back(vecf(arr("angles.f32", n[2])));
  c.rhs.push_back(vecf(arr("zshifts.f32", n[2])));
  c.rhs.push_back(i32(n[2]));
  c.rhs.push_back(i32(n[3]));
  c.rhs.push_back(i32(n[4]));
  c.rhs.push_back(i32(n[5]));
}

static std::string run(Call& c, int nlhs, mxArray** out) {
  try {
    mexFunction(nlhs, out, (int)c.rhs.size(), (const mxArray**)c.rhs.data());
    return "ok";
  } catch (const MexError& e) {
    return e.id + " " + e.what();
  }
}

int main(int argc, char** argv) {
  std::string mode = argc > 1 ? argv[1] : "errors";
  mxArray* out[2] = {nullptr, nullptr};
  if (mode == "errors") {
    {  // no action
      Call c;
      c.rhs.push_back(i32(3));
      std::cout << "noaction " << run(c, 1, out) << "\n";
    }
    {  // wrong arity for Proj
      Call c;
      geometry(c, "Proj", "");
      std::cout << "arity " << run(c, 1, out) << "\n";
    }
    {  // wrong class: image as double
      Call c;
      geometry(c, "Proj", "");
      c.rhs.push_back(f64(0));
      c.rhs.push_back(f32(1));
      c.rhs.push_back(f32(1));
      c.rhs.push_back(veci8({1, 1, 1, 1}));
      c.rhs.push_back(f64(0));
      c.rhs.push_back(i32(0));
      std::cout << "class " << run(c, 1, out) << "\n";
    }
    {  // unknown action
      Call c;
      c.rhs.push_back(mxCreateString("Spin"));
      std::cout << "action " << run(c, 1, out) << "\n";
    }
    {  // projector mode without a matched adjoint (pseudo-DD back, bit 4 set)
      Call c;
      geometry(c, "Back", "");
      c.rhs.push_back(vecf(std::vector<float>(8, 0.f)));
      c.rhs.push_back(f32(1));
      c.rhs.push_back(f32(1));
      c.rhs.push_back(veci8({1, 1, 1, 1}));
      c.rhs.push_back(f64(0));
      c.rhs.push_back(i32(0));
      c.rhs.push_back(i32(1 | 16));
      std::cout << "mode " << run(c, 1, out) << "\n";
    }
    return 0;
  }
  const std::string dir = argv[2];
  auto mask = readv<unsigned char>(dir + "/mask.u8");
  auto sc = readv<float>(dir + "/scalars.f32");
  if (mode == "leak") {
    for (int k = 0; k < 64; ++k) {
      Call c;
      geometry(c, "OSSART", dir);
      c.rhs.push_back(vecf(std::vector<float>(readv<float>(dir + "/vol.f32").size(), 0.f)));
      c.rhs.push_back(vecf(readv<float>(dir + "/sino.f32")));
      c.rhs.push_back(f32(sc[6]));
      c.rhs.push_back(f32(sc[7]));
      c.rhs.push_back(veci8(mask));
      c.rhs.push_back(f64(0));
      c.rhs.push_back(i32(4));
      c.rhs.push_back(i32(1));
      c.rhs.push_back(i32(0));
      out[0] = out[1] = nullptr;
      std::jmp_buf jb;
      ctr_debug_fail_alloc_after(k);
      volatile int failed = 0;
      if (setjmp(jb) == 0) {
        mex_error_jump() = &jb;
        mexFunction(2, out, (int)c.rhs.size(), (const mxArray**)c.rhs.data());
      } else {
        failed = 1;
      }
      mex_error_jump() = nullptr;
      ctr_debug_fail_alloc_after(-1);
      std::cout << "leak k=" << k << " error=" << failed << " plans=" << g_plans.size()
                << " buffers=" << g_bufs.size()
                << (failed ? " " + mex_error_text() : std::string()) << "\n";
      if (out[0]) mxDestroyArray(out[0]);
      if (out[1]) mxDestroyArray(out[1]);
      if (!failed) return 0;
    }
    return 1;
  }
  {
    Call c;
    geometry(c, "Proj", dir);
    c.rhs.push_back(vecf(readv<float>(dir + "/vol.f32")));
    c.rhs.push_back(f32(sc[6]));
    c.rhs.push_back(f32(sc[7]));
    c.rhs.push_back(veci8(mask));
    c.rhs.push_back(f64(0));  // gpu_id as dd3.m passes it (a double)
    c.rhs.push_back(i32(0));
    std::string r = run(c, 1, out);
    std::cout << "Proj " << r << "\n";
    if (r != "ok") return 1;
    writev(dir + "/proj_out.f32", mxGetData(out[0]), mxGetNumberOfElements(out[0]) * 4);
    mxDestroyArray(out[0]);
  }
  {
    Call c;
    geometry(c, "Back", dir);
    c.rhs.push_back(vecf(readv<float>(dir + "/sino.f32")));
    c.rhs.push_back(f32(sc[6]));
    c.rhs.push_back(f32(sc[7]));
    c.rhs.push_back(veci8(mask));
    c.rhs.push_back(f64(0));
    c.rhs.push_back(i32(0));
    c.rhs.push_back(i32(0));
    std::string r = run(c, 1, out);
    std::cout << "Back " << r << "\n";
    if (r != "ok") return 1;
    writev(dir + "/back_out.f32", mxGetData(out[0]), mxGetNumberOfElements(out[0]) * 4);
    mxDestroyArray(out[0]);
  }
  {
    Call c;
    geometry(c, "OSSART", dir);
    c.rhs.push_back(vecf(std::vector<float>(readv<float>(dir + "/vol.f32").size(), 0.f)));
    c.rhs.push_back(vecf(readv<float>(dir + "/sino.f32")));
    c.rhs.push_back(f32(sc[6]));
    c.rhs.push_back(f32(sc[7]));
    c.rhs.push_back(veci8(mask));
    c.rhs.push_back(f64(0));
    c.rhs.push_back(i32(4));  // osNum
    c.rhs.push_back(i32(2));  // iterNum
    c.rhs.push_back(i32(0));  // reference back projector
    std::string r = run(c, 2, out);
    std::cout << "OSSART " << r << "\n";
    if (r != "ok") return 1;
    writev(dir + "/sart_out.f32", mxGetData(out[0]), mxGetNumberOfElements(out[0]) * 4);
    std::cout << "resid " << mxGetPr(out[1])[0] << " " << mxGetPr(out[1])[1] << "\n";
    mxDestroyArray(out[0]);
    mxDestroyArray(out[1]);
  }
  {  // the argument list conjgrad.m passes
    Call c;
    geometry(c, "CG", dir);
    c.rhs.push_back(vecf(std::vector<float>(readv<float>(dir + "/vol.f32").size(), 0.f)));
    c.rhs.push_back(vecf(readv<float>(dir + "/sino.f32")));
    c.rhs.push_back(f32(sc[6]));
    c.rhs.push_back(f32(sc[7]));
    c.rhs.push_back(veci8(mask));
    c.rhs.push_back(f64(0));
    c.rhs.push_back(i32(3));  // iterNum
    c.rhs.push_back(i32(0));  // reference back projector
    std::string r = run(c, 1, out);
    std::cout << "CG " << r << "\n";
    if (r != "ok") return 1;
    writev(dir + "/cg_out.f32", mxGetData(out[0]), mxGetNumberOfElements(out[0]) * 4);
    mxDestroyArray(out[0]);
  }
  std::cout << "live plans=" << g_plans.size() << " buffers=" << g_bufs.size() << "\n";
  return g_plans.size() == 0 && g_bufs.size() == 0 ? 0 : 1;
}
