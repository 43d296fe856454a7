// Internal interface of the warm-start network kernels (unet.hip) used by the host
// driver (unet.cpp).  BF16 tensors are uint16_t arrays; see unet.hip for layouts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "ctr_internal.hpp"

namespace ctr {
namespace unet {

struct Grid3 {
  int Y, X, Z;
};

int conv3(const uint16_t* in, int cs_in, int Cin, const uint16_t* W, const float* bias,
          uint16_t* out, int cs_out, int Cout, bool relu, Grid3 g, hipStream_t st);
int wgrad_splits(Grid3 g);
int wgrad(const uint16_t* X, int cs_x, int Cin, const uint16_t* dY, int cs_dy, int Cout, Grid3 g,
          float* part, float* bpart, float* dW, float* db, hipStream_t st);
int convert_weights(const float* Wm, int Cin, int Cout, uint16_t* Wf, uint16_t* Wb,
                    hipStream_t st);
int pack_input(const float* x, const float* d, float s0, float s1, long long V, uint16_t* out,
               hipStream_t st);
int pool(const uint16_t* in, int cs_in, int C, Grid3 gf, uint16_t* out, int cs_out,
         hipStream_t st);
int pool_bwd(const uint16_t* dout, int cs_out, int C, Grid3 gf, uint16_t* din, int cs_in,
             hipStream_t st);
int upsample(const uint16_t* in, int cs_in, int C, Grid3 gf, uint16_t* out, int cs_out,
             hipStream_t st);
int upsample_bwd(const uint16_t* dout, int cs_out, int C, Grid3 gf, uint16_t* din, int cs_in,
                 hipStream_t st);
int relu_bwd(const uint16_t* y, int cs_y, const uint16_t* g, int cs_g, int C, long long V,
             uint16_t* dst, int cs_d, hipStream_t st);
int head(const uint16_t* a, const float* wb, long long V, float scale, float* h, hipStream_t st);
int mse(const float* h, const float* t, float tscale, long long V, float* dh, double* part,
        int nblocks, hipStream_t st);
int head_bwd(const uint16_t* a, const float* wb, const float* dh, long long V, uint16_t* da,
             double* part, int nblocks, float* dwb, hipStream_t st);
int adam(float* p, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
         float eps, int step, hipStream_t st);

}  // namespace unet
}  // namespace ctr
