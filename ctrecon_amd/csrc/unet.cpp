// Host driver of the learned warm-start network (the C ABI ctr_unet_*): layer table,
// parameter layout, activation / gradient buffers, forward, backward and Adam.
//
// Network (3-level U-Net, 3x3x3 convolutions, ReLU, 2x2x2 average pooling, nearest
// upsampling, skip connections by channel concatenation, a 1x1x1 head):
//   level 1 (full grid):  e1a 8->16, e1b 16->16 (skip s1)      ... d1a 48->16, d1b 16->16, head 16->1
//   level 2 (1/2):        e2a 16->32, e2b 32->32 (skip s2)     ... d2a 96->32, d2b 32->32
//   level 3 (1/4):        e3a 32->64, e3b 64->64
// Input channels: (x * s_x, delta * s_d, 0 x 6); output update = s_out * head.
// Parameters (FP32, in this order): for each conv layer W[27][Cout][Cin] then b[Cout],
// then the head w[16], b.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "unet_internal.hpp"

using namespace ctr;
using namespace ctr::unet;

namespace {

struct LayerSpec {
  int cin, cout;
};
constexpr int kNL = 10;
constexpr LayerSpec kL[kNL] = {{8, 16},  {16, 16}, {16, 32}, {32, 32}, {32, 64},
                               {64, 64}, {96, 32}, {32, 32}, {48, 16}, {16, 16}};

struct Layout {
  size_t woff[kNL], boff[kNL], head, n;
  Layout() {
    size_t o = 0;
    for (int l = 0; l < kNL; ++l) {
      woff[l] = o;
      o += (size_t)27 * kL[l].cin * kL[l].cout;
      boff[l] = o;
      o += kL[l].cout;
    }
    head = o;
    n = o + 17;
  }
};
const Layout kLay;

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  int ensure(size_t b) {
    if (bytes >= b) return CTR_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (dev_malloc(&p, b) != hipSuccess) {
      p = nullptr;
      return fail(CTR_ERR_NOMEM, "network buffer");
    }
    bytes = b;
    return CTR_OK;
  }
  uint16_t* h() const { return (uint16_t*)p; }
  float* f() const { return (float*)p; }
  double* d() const { return (double*)p; }
};

}  // namespace

struct ctr_unet {
  int device = 0;
  float s_x = 1.f, s_d = 1.f, s_out = 1.f;
  int step = 0;
  bool wdirty = true;
  Buf params, grads, m, v;
  Buf wf[kNL], wb[kNL];
  // activations (level 1 / 2 / 3 grids)
  Buf in0, a0, cat1, p1, a2, cat2, p2, a4, b3, a6, d2, a8, d1, h;
  // gradients
  Buf gA1, gB1, gC1, gA2, gB2, gC2, gB3, gC3, gP3, gP2, dh;
  Buf wpart, bpart, dpart, loss;
};

namespace {

int refresh_weights(ctr_unet& N, hipStream_t st) {
  if (!N.wdirty) return CTR_OK;
  for (int l = 0; l < kNL; ++l) {
    int rc = convert_weights(N.params.f() + kLay.woff[l], kL[l].cin, kL[l].cout, N.wf[l].h(),
                             N.wb[l].h(), st);
    if (rc) return rc;
  }
  N.wdirty = false;
  return CTR_OK;
}

int check_grid(int Y, int X, int Z) {
  if (Y < 4 || X < 4 || Z < 32 || Y % 4 || X % 4 || Z % 32)
    return fail(CTR_ERR_ARG, "network grid: Y and X must be multiples of 4 and Z of 32");
  return CTR_OK;
}

int ensure_buffers(ctr_unet& N, long long V, bool train) {
  const size_t V1 = (size_t)V, V2 = V1 / 8, V3 = V1 / 64;
  const size_t b = 2;  // bytes per BF16
  int rc;
  struct Req {
    Buf* buf;
    size_t bytes;
  } fwd[] = {{&N.in0, V1 * 8 * b},  {&N.a0, V1 * 16 * b}, {&N.cat1, V1 * 48 * b},
             {&N.p1, V2 * 16 * b},  {&N.a2, V2 * 32 * b}, {&N.cat2, V2 * 96 * b},
             {&N.p2, V3 * 32 * b},  {&N.a4, V3 * 64 * b}, {&N.b3, V3 * 64 * b},
             {&N.a6, V2 * 32 * b},  {&N.d2, V2 * 32 * b}, {&N.a8, V1 * 16 * b},
             {&N.d1, V1 * 16 * b},  {&N.h, V1 * 4}};
  for (auto& r : fwd)
    if ((rc = r.buf->ensure(r.bytes))) return rc;
  if (!train) return CTR_OK;
  const int nsplit = 64;
  size_t wmax = 0;
  for (int l = 0; l < kNL; ++l) wmax = std::max(wmax, (size_t)27 * kL[l].cin * kL[l].cout);
  Req bwd[] = {{&N.gA1, V1 * 48 * b}, {&N.gB1, V1 * 16 * b},  {&N.gC1, V1 * 16 * b},
               {&N.gA2, V2 * 96 * b}, {&N.gB2, V2 * 32 * b},  {&N.gC2, V2 * 32 * b},
               {&N.gB3, V3 * 64 * b}, {&N.gC3, V3 * 64 * b},  {&N.gP3, V3 * 32 * b},
               {&N.gP2, V2 * 16 * b}, {&N.dh, V1 * 4},         {&N.wpart, (size_t)nsplit * wmax * 4},
               {&N.bpart, (size_t)nsplit * 64 * 4}, {&N.dpart, (size_t)1024 * 17 * 8}, {&N.loss, 8}};
  for (auto& r : bwd)
    if ((rc = r.buf->ensure(r.bytes))) return rc;
  return CTR_OK;
}

#define TRY(x)            \
  do {                    \
    int rc_ = (x);        \
    if (rc_) return rc_;  \
  } while (0)

int forward(ctr_unet& N, const float* x, const float* d, Grid3 g1, float head_scale,
            hipStream_t st) {
  const long long V1 = (long long)g1.Y * g1.X * g1.Z;
  const Grid3 g2{g1.Y / 2, g1.X / 2, g1.Z / 2}, g3{g1.Y / 4, g1.X / 4, g1.Z / 4};
  const float* P = N.params.f();
  auto B = [&](int l) { return P + kLay.boff[l]; };
  TRY(refresh_weights(N, st));
  TRY(pack_input(x, d, N.s_x, N.s_d, V1, N.in0.h(), st));
  TRY(conv3(N.in0.h(), 8, 8, N.wf[0].h(), B(0), N.a0.h(), 16, 16, true, g1, st));
  TRY(conv3(N.a0.h(), 16, 16, N.wf[1].h(), B(1), N.cat1.h(), 48, 16, true, g1, st));  // s1
  TRY(pool(N.cat1.h(), 48, 16, g1, N.p1.h(), 16, st));
  TRY(conv3(N.p1.h(), 16, 16, N.wf[2].h(), B(2), N.a2.h(), 32, 32, true, g2, st));
  TRY(conv3(N.a2.h(), 32, 32, N.wf[3].h(), B(3), N.cat2.h(), 96, 32, true, g2, st));  // s2
  TRY(pool(N.cat2.h(), 96, 32, g2, N.p2.h(), 32, st));
  TRY(conv3(N.p2.h(), 32, 32, N.wf[4].h(), B(4), N.a4.h(), 64, 64, true, g3, st));
  TRY(conv3(N.a4.h(), 64, 64, N.wf[5].h(), B(5), N.b3.h(), 64, 64, true, g3, st));
  TRY(upsample(N.b3.h(), 64, 64, g2, N.cat2.h() + 32, 96, st));
  TRY(conv3(N.cat2.h(), 96, 96, N.wf[6].h(), B(6), N.a6.h(), 32, 32, true, g2, st));
  TRY(conv3(N.a6.h(), 32, 32, N.wf[7].h(), B(7), N.d2.h(), 32, 32, true, g2, st));
  TRY(upsample(N.d2.h(), 32, 32, g1, N.cat1.h() + 16, 48, st));
  TRY(conv3(N.cat1.h(), 48, 48, N.wf[8].h(), B(8), N.a8.h(), 16, 16, true, g1, st));
  TRY(conv3(N.a8.h(), 16, 16, N.wf[9].h(), B(9), N.d1.h(), 16, 16, true, g1, st));
  return head(N.d1.h(), P + kLay.head, V1, head_scale, N.h.f(), st);
}

// Loss and gradients of the last forward pass (head scale 1) against target t.
int backward(ctr_unet& N, const float* t, Grid3 g1, hipStream_t st, double* loss) {
  const long long V1 = (long long)g1.Y * g1.X * g1.Z, V2 = V1 / 8, V3 = V1 / 64;
  const Grid3 g2{g1.Y / 2, g1.X / 2, g1.Z / 2}, g3{g1.Y / 4, g1.X / 4, g1.Z / 4};
  const float* P = N.params.f();
  float* G = N.grads.f();
  auto GW = [&](int l) { return G + kLay.woff[l]; };
  auto GB = [&](int l) { return G + kLay.boff[l]; };
  float* wp = N.wpart.f();
  float* bp = N.bpart.f();
  long long nb = (V1 + 255) / 256;
  if (nb > 1024) nb = 1024;
  TRY(mse(N.h.f(), t, 1.f / N.s_out, V1, N.dh.f(), N.dpart.d(), (int)nb, st));
  if (loss) {
    std::vector<double> h(nb);
    if (hipMemcpyAsync(h.data(), N.dpart.p, nb * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(CTR_ERR_HIP, "loss readback");
    double s = 0;
    for (double x : h) s += x;
    *loss = s / (2.0 * (double)V1);
  }
  // head (with the ReLU of d1b)
  TRY(head_bwd(N.d1.h(), P + kLay.head, N.dh.f(), V1, N.gB1.h(), N.dpart.d(), (int)nb,
               G + kLay.head, st));
  // d1b: input a8
  TRY(wgrad(N.a8.h(), 16, 16, N.gB1.h(), 16, 16, g1, wp, bp, GW(9), GB(9), st));
  TRY(conv3(N.gB1.h(), 16, 16, N.wb[9].h(), nullptr, N.gC1.h(), 16, 16, false, g1, st));
  TRY(relu_bwd(N.a8.h(), 16, N.gC1.h(), 16, 16, V1, N.gC1.h(), 16, st));
  // d1a: input cat1
  TRY(wgrad(N.cat1.h(), 48, 48, N.gC1.h(), 16, 16, g1, wp, bp, GW(8), GB(8), st));
  TRY(conv3(N.gC1.h(), 16, 16, N.wb[8].h(), nullptr, N.gA1.h(), 48, 48, false, g1, st));
  TRY(upsample_bwd(N.gA1.h() + 16, 48, 32, g1, N.gB2.h(), 32, st));
  TRY(relu_bwd(N.d2.h(), 32, N.gB2.h(), 32, 32, V2, N.gB2.h(), 32, st));
  // d2b: input a6
  TRY(wgrad(N.a6.h(), 32, 32, N.gB2.h(), 32, 32, g2, wp, bp, GW(7), GB(7), st));
  TRY(conv3(N.gB2.h(), 32, 32, N.wb[7].h(), nullptr, N.gC2.h(), 32, 32, false, g2, st));
  TRY(relu_bwd(N.a6.h(), 32, N.gC2.h(), 32, 32, V2, N.gC2.h(), 32, st));
  // d2a: input cat2
  TRY(wgrad(N.cat2.h(), 96, 96, N.gC2.h(), 32, 32, g2, wp, bp, GW(6), GB(6), st));
  TRY(conv3(N.gC2.h(), 32, 32, N.wb[6].h(), nullptr, N.gA2.h(), 96, 96, false, g2, st));
  TRY(upsample_bwd(N.gA2.h() + 32, 96, 64, g2, N.gB3.h(), 64, st));
  TRY(relu_bwd(N.b3.h(), 64, N.gB3.h(), 64, 64, V3, N.gB3.h(), 64, st));
  // e3b: input a4
  TRY(wgrad(N.a4.h(), 64, 64, N.gB3.h(), 64, 64, g3, wp, bp, GW(5), GB(5), st));
  TRY(conv3(N.gB3.h(), 64, 64, N.wb[5].h(), nullptr, N.gC3.h(), 64, 64, false, g3, st));
  TRY(relu_bwd(N.a4.h(), 64, N.gC3.h(), 64, 64, V3, N.gC3.h(), 64, st));
  // e3a: input p2
  TRY(wgrad(N.p2.h(), 32, 32, N.gC3.h(), 64, 64, g3, wp, bp, GW(4), GB(4), st));
  TRY(conv3(N.gC3.h(), 64, 64, N.wb[4].h(), nullptr, N.gP3.h(), 32, 32, false, g3, st));
  TRY(pool_bwd(N.gP3.h(), 32, 32, g2, N.gA2.h(), 96, st));  // d s2 += pool adjoint
  TRY(relu_bwd(N.cat2.h(), 96, N.gA2.h(), 96, 32, V2, N.gA2.h(), 96, st));
  // e2b: input a2, gradient = gA2[:, :32]
  TRY(wgrad(N.a2.h(), 32, 32, N.gA2.h(), 96, 32, g2, wp, bp, GW(3), GB(3), st));
  TRY(conv3(N.gA2.h(), 96, 32, N.wb[3].h(), nullptr, N.gB2.h(), 32, 32, false, g2, st));
  TRY(relu_bwd(N.a2.h(), 32, N.gB2.h(), 32, 32, V2, N.gB2.h(), 32, st));
  // e2a: input p1
  TRY(wgrad(N.p1.h(), 16, 16, N.gB2.h(), 32, 32, g2, wp, bp, GW(2), GB(2), st));
  TRY(conv3(N.gB2.h(), 32, 32, N.wb[2].h(), nullptr, N.gP2.h(), 16, 16, false, g2, st));
  TRY(pool_bwd(N.gP2.h(), 16, 16, g1, N.gA1.h(), 48, st));  // d s1 += pool adjoint
  TRY(relu_bwd(N.cat1.h(), 48, N.gA1.h(), 48, 16, V1, N.gA1.h(), 48, st));
  // e1b: input a0, gradient = gA1[:, :16]
  TRY(wgrad(N.a0.h(), 16, 16, N.gA1.h(), 48, 16, g1, wp, bp, GW(1), GB(1), st));
  TRY(conv3(N.gA1.h(), 48, 16, N.wb[1].h(), nullptr, N.gB1.h(), 16, 16, false, g1, st));
  TRY(relu_bwd(N.a0.h(), 16, N.gB1.h(), 16, 16, V1, N.gB1.h(), 16, st));
  // e1a: input in0
  return wgrad(N.in0.h(), 8, 8, N.gB1.h(), 16, 16, g1, wp, bp, GW(0), GB(0), st);
}

}  // namespace

extern "C" {

long long ctr_unet_num_params(void) { return (long long)kLay.n; }

int ctr_unet_create(int device, ctr_unet** out) {
  if (!out) return fail(CTR_ERR_ARG, "out is NULL");
  *out = nullptr;
  int n = 0;
  CTR_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(CTR_ERR_ARG, "device not present");
  CTR_HIP(hipSetDevice(device));
  ctr_unet* N = new ctr_unet();
  N->device = device;
  int rc;
  const size_t pb = kLay.n * sizeof(float);
  if ((rc = N->params.ensure(pb)) || (rc = N->grads.ensure(pb)) || (rc = N->m.ensure(pb)) ||
      (rc = N->v.ensure(pb))) {
    delete N;
    return rc;
  }
  for (int l = 0; l < kNL; ++l) {
    const size_t wb = (size_t)27 * kL[l].cin * kL[l].cout * 2;
    if ((rc = N->wf[l].ensure(wb)) || (rc = N->wb[l].ensure(wb))) {
      delete N;
      return rc;
    }
  }
  if (hipMemset(N->params.p, 0, pb) != hipSuccess || hipMemset(N->m.p, 0, pb) != hipSuccess ||
      hipMemset(N->v.p, 0, pb) != hipSuccess || hipMemset(N->grads.p, 0, pb) != hipSuccess) {
    delete N;
    return fail(CTR_ERR_HIP, "network parameter initialisation");
  }
  *out = N;
  return CTR_OK;
}

int ctr_unet_destroy(ctr_unet* net) {
  if (net) {
    (void)hipSetDevice(net->device);
    delete net;
  }
  return CTR_OK;
}

int ctr_unet_set_params(ctr_unet* net, const float* host, long long n) {
  if (!net || !host) return fail(CTR_ERR_ARG, "NULL argument");
  if (n != (long long)kLay.n) return fail(CTR_ERR_ARG, "parameter count mismatch");
  CTR_HIP(hipSetDevice(net->device));
  const size_t pb = kLay.n * sizeof(float);
  CTR_HIP(hipMemcpy(net->params.p, host, pb, hipMemcpyHostToDevice));
  CTR_HIP(hipMemset(net->m.p, 0, pb));
  CTR_HIP(hipMemset(net->v.p, 0, pb));
  net->step = 0;
  net->wdirty = true;
  return CTR_OK;
}

int ctr_unet_get_params(ctr_unet* net, float* host, long long n) {
  if (!net || !host) return fail(CTR_ERR_ARG, "NULL argument");
  if (n != (long long)kLay.n) return fail(CTR_ERR_ARG, "parameter count mismatch");
  CTR_HIP(hipSetDevice(net->device));
  CTR_HIP(hipMemcpy(host, net->params.p, kLay.n * sizeof(float), hipMemcpyDeviceToHost));
  return CTR_OK;
}

int ctr_unet_set_scales(ctr_unet* net, float s_x, float s_d, float s_out) {
  if (!net) return fail(CTR_ERR_ARG, "NULL network");
  if (!(s_out != 0.f) || !std::isfinite(s_x) || !std::isfinite(s_d) || !std::isfinite(s_out))
    return fail(CTR_ERR_ARG, "scales must be finite and s_out non-zero");
  net->s_x = s_x;
  net->s_d = s_d;
  net->s_out = s_out;
  return CTR_OK;
}

int ctr_unet_forward(ctr_unet* net, const float* d_x, const float* d_delta, float* d_update,
                     int Y, int X, int Z, void* stream) {
  if (!net || !d_x || !d_delta || !d_update) return fail(CTR_ERR_ARG, "NULL argument");
  int rc = check_grid(Y, X, Z);
  if (rc) return rc;
  CTR_HIP(hipSetDevice(net->device));
  hipStream_t st = (hipStream_t)stream;
  const long long V = (long long)Y * X * Z;
  if ((rc = ensure_buffers(*net, V, false))) return rc;
  if ((rc = forward(*net, d_x, d_delta, Grid3{Y, X, Z}, net->s_out, st))) return rc;
  CTR_HIP(hipMemcpyAsync(d_update, net->h.p, V * sizeof(float), hipMemcpyDeviceToDevice, st));
  return CTR_OK;
}

int ctr_unet_loss_grad(ctr_unet* net, const float* d_x, const float* d_delta,
                       const float* d_target, int Y, int X, int Z, float* host_grads,
                       double* loss) {
  if (!net || !d_x || !d_delta || !d_target) return fail(CTR_ERR_ARG, "NULL argument");
  int rc = check_grid(Y, X, Z);
  if (rc) return rc;
  CTR_HIP(hipSetDevice(net->device));
  const long long V = (long long)Y * X * Z;
  if ((rc = ensure_buffers(*net, V, true))) return rc;
  const Grid3 g{Y, X, Z};
  if ((rc = forward(*net, d_x, d_delta, g, 1.f, nullptr))) return rc;
  if ((rc = backward(*net, d_target, g, nullptr, loss))) return rc;
  if (host_grads)
    CTR_HIP(hipMemcpy(host_grads, net->grads.p, kLay.n * sizeof(float), hipMemcpyDeviceToHost));
  CTR_HIP(hipDeviceSynchronize());
  return CTR_OK;
}

int ctr_unet_train_step(ctr_unet* net, const float* d_x, const float* d_delta,
                        const float* d_target, int Y, int X, int Z, float lr, double* loss) {
  if (!net || !d_x || !d_delta || !d_target) return fail(CTR_ERR_ARG, "NULL argument");
  if (!(lr > 0.f)) return fail(CTR_ERR_ARG, "lr must be positive");
  int rc = check_grid(Y, X, Z);
  if (rc) return rc;
  CTR_HIP(hipSetDevice(net->device));
  const long long V = (long long)Y * X * Z;
  if ((rc = ensure_buffers(*net, V, true))) return rc;
  const Grid3 g{Y, X, Z};
  if ((rc = forward(*net, d_x, d_delta, g, 1.f, nullptr))) return rc;
  if ((rc = backward(*net, d_target, g, nullptr, loss))) return rc;
  net->step += 1;
  if ((rc = adam(net->params.f(), net->grads.f(), net->m.f(), net->v.f(), kLay.n, lr, 0.9f,
                 0.999f, 1e-8f, net->step, nullptr)))
    return rc;
  net->wdirty = true;
  return CTR_OK;
}

}  // extern "C"
