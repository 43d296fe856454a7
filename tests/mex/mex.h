// Minimal stand-in for the MATLAB C matrix / MEX API, enough to compile and drive
// the DD3_GPU gateway without MATLAB.  TEST INFRASTRUCTURE ONLY.
// mexErrMsgIdAndTxt throws MexError so a driver can check error identifiers; when a
// driver sets mex_error_jump(), it longjmps there instead, leaving the gateway the way
// MATLAB does (no C++ destructors run).
#pragma once
#include <csetjmp>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

typedef size_t mwSize;
enum mxClassID { mxSINGLE_CLASS, mxDOUBLE_CLASS, mxINT32_CLASS, mxINT8_CLASS, mxUINT8_CLASS,
                 mxCHAR_CLASS };
enum mxComplexity { mxREAL };

struct mxArray {
  mxClassID cls;
  std::vector<mwSize> dims;
  std::vector<unsigned char> data;
  std::string str;
};

struct MexError : std::runtime_error {
  std::string id;
  MexError(const std::string& i, const std::string& m) : std::runtime_error(m), id(i) {}
};

inline size_t mx_elsize(mxClassID c) {
  switch (c) {
    case mxSINGLE_CLASS: case mxINT32_CLASS: return 4;
    case mxDOUBLE_CLASS: return 8;
    default: return 1;
  }
}
inline size_t mxGetNumberOfElements(const mxArray* a) {
  size_t n = 1;
  for (auto d : a->dims) n *= d;
  return a->cls == mxCHAR_CLASS ? a->str.size() : n;
}
inline void* mxGetData(const mxArray* a) { return (void*)a->data.data(); }
inline double* mxGetPr(const mxArray* a) { return (double*)a->data.data(); }
inline bool mxIsSingle(const mxArray* a) { return a->cls == mxSINGLE_CLASS; }
inline bool mxIsDouble(const mxArray* a) { return a->cls == mxDOUBLE_CLASS; }
inline bool mxIsInt32(const mxArray* a) { return a->cls == mxINT32_CLASS; }
inline bool mxIsInt8(const mxArray* a) { return a->cls == mxINT8_CLASS; }
inline bool mxIsUint8(const mxArray* a) { return a->cls == mxUINT8_CLASS; }
inline bool mxIsChar(const mxArray* a) { return a->cls == mxCHAR_CLASS; }
inline double mxGetScalar(const mxArray* a) {
  switch (a->cls) {
    case mxDOUBLE_CLASS: return *(const double*)a->data.data();
    case mxSINGLE_CLASS: return *(const float*)a->data.data();
    case mxINT32_CLASS: return *(const int32_t*)a->data.data();
    default: return a->data.empty() ? 0.0 : (double)a->data[0];
  }
}
inline int mxGetString(const mxArray* a, char* buf, mwSize len) {
  if (a->cls != mxCHAR_CLASS || a->str.size() + 1 > len) return 1;
  std::memcpy(buf, a->str.c_str(), a->str.size() + 1);
  return 0;
}
inline mxArray* mxCreateNumericArray(mwSize nd, const mwSize* dims, mxClassID c, mxComplexity) {
  mxArray* a = new mxArray{c, std::vector<mwSize>(dims, dims + nd), {}, {}};
  a->data.assign(mxGetNumberOfElements(a) * mx_elsize(c), 0);
  return a;
}
inline mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity r) {
  const mwSize d[2] = {m, n};
  return mxCreateNumericArray(2, d, mxDOUBLE_CLASS, r);
}
inline mxArray* mxCreateString(const char* s) {
  mxArray* a = new mxArray{mxCHAR_CLASS, {1, std::strlen(s)}, {}, s};
  return a;
}
inline void mxDestroyArray(mxArray* a) { delete a; }
inline std::jmp_buf*& mex_error_jump() {
  static std::jmp_buf* j = nullptr;
  return j;
}
inline std::string& mex_error_text() {
  static std::string s;
  return s;
}
[[noreturn]] inline void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (mex_error_jump()) {
    mex_error_text() = std::string(id) + " " + buf;
    std::longjmp(*mex_error_jump(), 1);
  }
  throw MexError(id, buf);
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
