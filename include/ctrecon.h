/*
 * ctrecon.h -- C ABI of the MI355X DD3 projector library (libctrecon_amd.so).
 *
 * Drop-in boundary for the reference's DD3 GPU projector pair and the OS-SART /
 * CG loops that drive it (LittleRabbitRun/CTRecon).  Plain pointers and sizes
 * only; every entry point returns a status code (CTR_OK == 0) and never exits
 * the process; ctr_last_error() gives a thread-local message.
 *
 * Data layouts (as in the reference GPU code):
 *   volume    vol[(y*XN + x)*ZN + z]         (iterativeRecon/DD3_GPU.cu:44,1473)
 *   sinogram  prj[(p*DNU + u)*DNV + v]       (iterativeRecon/DD3_GPU.cu:951)
 *   mask      mask[y*XN + x], unsigned bytes (iterativeRecon/DD3_GPU.cu:1004-1011,1417)
 * Geometry arguments are those of the GPU path after dd3.m's axis swap
 * (angles already +pi/2, iterativeRecon/dd3.m:247-250,305-307).
 */
#ifndef CTRECON_H
#define CTRECON_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define CTR_OK 0
#define CTR_ERR_ARG 1         /* invalid argument (null pointer, bad size, bad mode) */
#define CTR_ERR_HIP 2         /* HIP runtime error */
#define CTR_ERR_UNSUPPORTED 3 /* projector mode or geometry this build does not implement */
#define CTR_ERR_NOMEM 4       /* device allocation failed */
#define CTR_ERR_NUMERIC 5     /* non-finite input where a finite one is required */

/* ---- projector modes (prjMode argument of the reference, iterativeRecon/dd3.m:233-246) ---- */
#define CTR_MODE_BRANCHLESS 0   /* distance-driven, branchless (the reference's mode 0 / default) */
#define CTR_MODE_PSEUDO_DD 1    /* forward and back: pseudo distance-driven (DD3_GPU.cu:1153, 1787) */
#define CTR_MODE_FP64_SAT 2     /* forward: FP64 branchless (DD3_GPU.cu:1161); back: pseudo-DD (:1794) */
#define CTR_MODE_ZLINE 3        /* forward: pseudo-DD (:1168); back: z-line branchless (:1801) */
/* Back projector selector (build addition, no reference mode number): the exact
 * transpose of the mode-0 forward projector, used for adjoint tests and CG. */
#define CTR_BP_REFERENCE 0
#define CTR_BP_MATCHED 1

/* Projector method of a plan (build addition).  DIRECT integrates each footprint
 * over the cells it overlaps; SAT samples compensated (double-float) integral
 * images with clamped bilinear weights, the reference's formulation
 * (iterativeRecon/DD3_GPU.cu:134-197, 844-990, 1284-1323, 1397-1475).  Both compute
 * the same operator; the matched adjoint is always DIRECT. */
#define CTR_METHOD_DIRECT 0
#define CTR_METHOD_SAT 1
/* The reference's pseudo distance-driven pair (prjMode 1 / 3 forward, 1 / 2 back;
 * iterativeRecon/DD3_GPU.cu:673-834, 1478-1528): one ray per detector cell sampled
 * bilinearly on the slice centre planes, and a voxel-driven bilinear sinogram sample per
 * view.  A different operator from the two above; FP32 operands only, reference back
 * projector only (no matched adjoint), not available to the OS-SART / CG drivers. */
#define CTR_METHOD_PSEUDO_DD 2
/* The reference's prjMode 3 pair: the pseudo-DD forward projector and the z-line
 * back projector (iterativeRecon/DD3_GPU.cu:1545-1647), whose reference comment notes
 * that it needs ZN divisible by 16: it writes the slices of the whole 16-slice blocks
 * and leaves the rest 0, which is reproduced.  Same restrictions as PSEUDO_DD. */
#define CTR_METHOD_ZLINE 3

/* Storage of projector operands (build addition; the reference is FP32 throughout).
 * BF16 arrays are unsigned short bit patterns: the upper half of the FP32 value,
 * rounded to nearest even.  Every sum is accumulated in FP32 or wider. */
#define CTR_STORAGE_F32 0
#define CTR_STORAGE_BF16 1      /* volume and sinogram operands BF16 */
#define CTR_STORAGE_BF16_SINO 2 /* sinogram operands BF16, the iterate read in FP32 */

/*
 * Forward projection, host buffers in and out.
 * Replaces DD3Proj_gpu (iterativeRecon/DD3_GPU.cu:1132-1142), same parameter list and
 * the same prjMode dispatch (:1144-1183): 1 and 3 the pseudo distance-driven projector,
 * 2 the FP64 branchless projector (computed here in FP32: the footprint model of mode 0
 * sampled half a cell off across the rays and along z, as the reference's point-sampled
 * integral images are), 0 and every other value the branchless distance-driven model.
 * Unlike the reference, hvol is not modified (the mask is applied on the device) and
 * the device is not reset.
 */
int ctr_dd3_proj(float x0, float y0, float z0, int DNU, int DNV, const float* xds,
                 const float* yds, const float* zds, float imgXCenter, float imgYCenter,
                 float imgZCenter, const float* hangs, const float* hzPos, int PN, int XN,
                 int YN, int ZN, const float* hvol, float* hprj, float dx, float dz,
                 const unsigned char* mask, int gpunum, int prjMode);

/*
 * Back projection, host buffers in and out.
 * Replaces DD3Back_gpu (iterativeRecon/DD3_GPU.cu:1766-1775), same parameter list
 * (squared is accepted and ignored, as in the reference, :1393) and the same prjMode
 * dispatch (:1778-1815): 1 and 2 the pseudo distance-driven back projector, 3 the z-line
 * branchless back projector, 0 and every other value the reference-arithmetic
 * branchless back projector; bit 4 (0 | 16) selects the matched adjoint of the mode-0
 * forward projector instead.  Voxels whose mask byte is not 1 are 0 (:1417, :1496).
 */
int ctr_dd3_back(float x0, float y0, float z0, int DNU, int DNV, const float* xds,
                 const float* yds, const float* zds, float imgXCenter, float imgYCenter,
                 float imgZCenter, const float* hangs, const float* hzPos, int PN, int XN,
                 int YN, int ZN, float* hvol, const float* hprj, float dx, float dz,
                 const unsigned char* mask, int gpunum, int squared, int prjMode);

/* ---- persistent, device-resident plan ---- */

/* The geometry of one scan: the non-data arguments of DD3Proj_gpu
 * (iterativeRecon/DD3_GPU.cu:1133-1142). */
typedef struct ctr_geometry {
  float x0, y0, z0;
  int DNU, DNV;
  const float* xds; /* [DNU] */
  const float* yds; /* [DNU] */
  const float* zds; /* [DNV], strictly increasing */
  float imgXCenter, imgYCenter, imgZCenter;
  const float* angles;  /* [PN] */
  const float* zshifts; /* [PN] */
  int PN;
  int XN, YN, ZN;
  float dx, dz;
  const unsigned char* mask; /* [XN*YN] or NULL for all ones */
} ctr_geometry;

typedef struct ctr_plan ctr_plan;

/* Build per-view tables on `device` (host FP64 arithmetic, stored FP32).
 * Replaces the per-call set-up of DD3_gpu_proj_branchless_sat2d / DD3_gpu_back
 * (iterativeRecon/DD3_GPU.cu:1012-1079, 1664-1686), done once per scan. */
int ctr_plan_create(const ctr_geometry* geom, int device, ctr_plan** out);
int ctr_plan_destroy(ctr_plan* plan);

/* Select the projector method (CTR_METHOD_DIRECT, the default, CTR_METHOD_SAT,
 * CTR_METHOD_PSEUDO_DD or CTR_METHOD_ZLINE) used by every later call on the plan.  The reference selects
 * the model per call with prjMode and always builds integral images for mode 0
 * (iterativeRecon/DD3_GPU.cu:1036, 1710). */
int ctr_plan_set_method(ctr_plan* plan, int method);

/* Storage the OS-SART and CG drivers use for their projector operands (build
 * addition, no reference counterpart).  CTR_STORAGE_BF16: the iterate keeps its FP32
 * master copy (d_x) and a dithered BF16 shadow that the forward projector reads; the
 * sinogram passed from the forward to the back projector is BF16.
 * CTR_STORAGE_BF16_SINO: only that sinogram is BF16 (the forward projector reads the
 * FP32 iterate).  In every mode the measured sinogram d_prj, the row / column sums, the
 * update and every reduction stay FP32 / FP64.  ctr_residual always projects in FP32. */
int ctr_plan_set_storage(ctr_plan* plan, int storage);

/* Forward-project device volume d_vol into device sinogram d_prj for the views
 * view_begin + k*view_stride, k < view_count; d_prj is [view_count][DNU][DNV].
 * The kernel of DD3_gpu_proj_branchless_sat2d_ker (iterativeRecon/DD3_GPU.cu:844-990).
 * stream is a hipStream_t (NULL: the null stream). Asynchronous. */
int ctr_plan_fp(ctr_plan* plan, const float* d_vol, float* d_prj, int view_begin,
                int view_stride, int view_count, void* stream);

/* Back-project device sinogram d_prj ([view_count][DNU][DNV], views as in
 * ctr_plan_fp) into device volume d_vol (overwritten).  bp_kind: CTR_BP_REFERENCE
 * (DD3_gpu_back_ker<_BRANCHLESS>, iterativeRecon/DD3_GPU.cu:1397-1475) or
 * CTR_BP_MATCHED (exact transpose of ctr_plan_fp).  Asynchronous. */
int ctr_plan_bp(ctr_plan* plan, const float* d_prj, float* d_vol, int view_begin,
                int view_stride, int view_count, int bp_kind, void* stream);

/* d_vol[y][x][z] *= mask[y*XN + x]: the reference's in-place masking of the volume
 * (iterativeRecon/DD3_GPU.cu:1004-1011) on a device volume.  Asynchronous. */
int ctr_plan_mask_volume(ctr_plan* plan, float* d_vol, void* stream);

/* ctr_plan_fp / ctr_plan_bp with BF16 storage (build addition): d_vol and d_prj of
 * the forward projector, and d_prj of the back projector, are BF16 arrays; the back
 * projector's output d_vol is FP32.  Asynchronous. */
int ctr_plan_fp_bf16(ctr_plan* plan, const unsigned short* d_vol, unsigned short* d_prj,
                     int view_begin, int view_stride, int view_count, void* stream);
int ctr_plan_bp_bf16(ctr_plan* plan, const unsigned short* d_prj, float* d_vol, int view_begin,
                     int view_stride, int view_count, int bp_kind, void* stream);

/* OS-SART, iterativeRecon/os_sart.m:1-44, on the device.  d_x (in/out) is the
 * volume iterate, d_prj the full measured sinogram [PN][DNU][DNV].  Subset s holds
 * views s, s+osNum, ...  resid (host, may be NULL) receives iterNum values: the
 * relative data residual ||b - A x||/||b|| accumulated (FP64) over each sweep.
 * Synchronous. */
int ctr_os_sart(ctr_plan* plan, float* d_x, const float* d_prj, int osNum, int iterNum,
                int bp_kind, double* resid);

/* Timing of ctr_os_sart_run (HIP events on the launch stream). */
typedef struct ctr_sart_stats {
  int sweeps;          /* sweeps run */
  int fp_launches;     /* fused forward-projection launches timed */
  int bp_launches;     /* back-projection (update) launches timed */
  double setup_ms;     /* row / column sums and their guarded inverses */
  double fp_ms;        /* summed duration of the forward-projection launches */
  double bp_ms;        /* summed duration of the back-projection launches */
} ctr_sart_stats;

/* ctr_os_sart with a stopping rule and timing (no counterpart in os_sart.m, which
 * has neither): runs at most maxIter sweeps, stops after the first sweep whose
 * residual is <= stop_resid (stop_resid <= 0: never).  resid / sweep_ms (host,
 * maxIter entries, may be NULL) receive the per-sweep residual and duration;
 * stats (may be NULL) the set-up time and per-kernel totals.  Synchronous. */
int ctr_os_sart_run(ctr_plan* plan, float* d_x, const float* d_prj, int osNum, int maxIter,
                    int bp_kind, double stop_resid, double* resid, double* sweep_ms,
                    ctr_sart_stats* stats);

/* OS-SART session (build addition): the subset row / column sums of one measured
 * sinogram (os_sart.m:14-18) computed once by ctr_sart_create, then any number of
 * ctr_sart_sweeps calls on an iterate that the caller may modify in between (the
 * learned warm start does).  The session uses the plan's method and storage at
 * creation; d_prj must stay valid until ctr_sart_destroy.  Each session owns its
 * reduction scratch; every session of a plan runs on the plan's stream, so calls on one
 * plan must come from one thread at a time. */
typedef struct ctr_sart ctr_sart;
int ctr_sart_create(ctr_plan* plan, const float* d_prj, int osNum, int bp_kind,
                    ctr_sart** out);
/* Sweeps as in ctr_os_sart_run (stats->setup_ms is the session's set-up time).  With
 * the direct method and the reference back projector, and stats == NULL, the first call
 * for an iterate d_x captures one sweep (3 x osNum launches on the plan's stream) into a
 * HIP graph and every sweep replays it; a new d_x is captured again.  With stats, the
 * per-launch timing events need plain launches. */
int ctr_sart_sweeps(ctr_sart* sart, float* d_x, int maxIter, double stop_resid, double* resid,
                    double* sweep_ms, ctr_sart_stats* stats);
/* Graph replay on (default) or plain launches for the session's later sweeps. */
int ctr_sart_set_graph(ctr_sart* sart, int enable);
/* Whether a sweep graph is held, and the host time its capture + instantiation took. */
int ctr_sart_graph_info(const ctr_sart* sart, int* captured, double* capture_ms);
double ctr_sart_setup_ms(const ctr_sart* sart);
int ctr_sart_destroy(ctr_sart* sart);

/* Number of masked voxels inside each view's cone (the voxels a view can see:
 * z-cells meeting [sz + (bz_0 - z0) l/S2D, sz + (bz_DNV - z0) l/S2D] at in-plane
 * source distance l), for views view_begin + k*view_stride.  Host computation;
 * the per-view volume term of the algorithmic byte model. */
int ctr_cone_voxels(ctr_plan* plan, int view_begin, int view_stride, int view_count,
                    long long* out);

/* CG, iterativeRecon/conjgrad.m:1-35, on the device: A = Back(Proj(.)),
 * b = Back(prj); dot products and step scalars in FP64.  resid (host, may be
 * NULL) receives ||r_k|| per iteration.  The CG vectors are kept in the plan: the first
 * call allocates them, later calls of the same storage allocate nothing.  Runs on the
 * plan's stream.  Synchronous. */
int ctr_conjgrad(ctr_plan* plan, float* d_x, const float* d_prj, int iterNum, int bp_kind,
                 double* resid);

/* Relative data residual ||A x - b|| / ||b|| over all views (FP64 sums). Synchronous. */
int ctr_residual(ctr_plan* plan, const float* d_x, const float* d_prj, double* out);

/* ctr_residual for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc); no counterpart in the
 * reference.  Same result as ctr_residual, bit for bit, and it calls no hipMalloc / hipFree:
 * its reduction scratch is a fixed 4 MiB device array of the library (one per GPU, held
 * until the process ends), enough for detectors of DNU * DNV up to 261888 rays when the
 * scan has 256 views or more (more rays with fewer views; CTR_ERR_ARG beyond).  Direct
 * method only (CTR_ERR_UNSUPPORTED otherwise).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (several threads or plans at
 * once) race and must be serialised by the caller.  Synchronous. */
int ctr_residual_noalloc(ctr_plan* plan, const float* d_x, const float* d_prj, double* out);

/* ---- batched OS-SART over several scans (build addition, no reference counterpart) ----
 * nscan scans of any sizes (one plan each, same device and storage, direct method),
 * their iterates d_x[i] and measured sinograms d_prj[i] (kept by the batch until
 * destroyed).  ctr_batch_create computes every scan's row / column sums and tunes the
 * launch shapes on the batch (persistent grid sizes, forward work mapping, z groups per
 * lane of the back projector, contiguous / interleaved / one-per-item work assignment,
 * chosen by timing; the choice is cached in a file keyed by the device and the batch's
 * geometry -- $CTRECON_TUNE_CACHE, else ~/.cache/ctrecon_amd/batch_tuning.txt -- and a
 * later batch with the same key takes it from there); each subset step then runs as one
 * forward, one reduction and one back-projection launch per back-projection class over all
 * scans, and a whole sweep is captured once into a HIP graph and replayed.  Reference back projector
 * only. */
typedef struct ctr_batch ctr_batch;
typedef struct ctr_batch_info {
  int fp_colw;        /* forward mapping: 1 = one wave per detector column, 0 = per ray */
  int fp_grid;        /* persistent workgroups of the forward launch (0: one per item) */
  int fp_strided;     /* 1: interleaved item ranges, 0: contiguous */
  int bp_zg;          /* back projector: z groups per lane 1..4, 0: each scan's ceil(ZN/64),
                         5: windowed (one wave per column over all z, sums in LDS) */
  int bp_grid;        /* persistent workgroups of the back-projection launch (0: one per item) */
  int bp_strided;     /* 1: interleaved item ranges, 0: contiguous */
  double setup_ms;    /* row / column sums of all scans */
  double tune_ms;     /* autotuning (host wall time) */
  double capture_ms;  /* graph capture + instantiation (host wall time; 0 before) */
  long long fp_items, bp_items;  /* work items per sweep */
  int from_cache;     /* 1: the launch shapes came from the tuning cache (tune_ms: lookup) */
} ctr_batch_info;
int ctr_batch_create(ctr_plan* const* plans, int nscan, float* const* d_x,
                     const float* const* d_prj, int osNum, int bp_kind, ctr_batch** out);
/* Up to maxIter sweeps; stops after the first sweep at which every scan's residual is
 * <= stop_resid (stop_resid <= 0: never).  resid (host, may be NULL) receives
 * [sweep][scan] residuals, sweep_ms the per-sweep GPU time, done the sweeps run. */
int ctr_batch_sweeps(ctr_batch* batch, int maxIter, double stop_resid, double* resid,
                     double* sweep_ms, int* done);
/* Graph replay on (default) or plain launches. */
int ctr_batch_set_graph(ctr_batch* batch, int enable);
/* Override the tuned launch shapes (tests and experiments). */
int ctr_batch_set_tuning(ctr_batch* batch, int fp_colw, int fp_grid, int fp_strided, int bp_zg,
                         int bp_grid, int bp_strided);
int ctr_batch_get_info(const ctr_batch* batch, ctr_batch_info* info);
int ctr_batch_destroy(ctr_batch* batch);

/* ---- learned warm start (build addition, no reference counterpart) ----
 * A 3-level 3-D U-Net (16/32/64 channels, 3x3x3 convolutions, ReLU, 2x2x2 average
 * pooling, nearest upsampling, concatenated skips, 1x1x1 head) run with BF16
 * activations and weights and FP32 accumulation on the matrix cores; FP32 master
 * parameters and Adam.  Input: the OS-SART iterate x and the last sweep's update
 * delta (FP32 volumes, vol[(y*X + x)*Z + z]); output: the predicted update
 * s_out * net(x * s_x, delta * s_d).  Grids: Y, X multiples of 4, Z a multiple of 32.
 * Parameters: for each of the 10 convolutions W[27][Cout][Cin] then b[Cout], then the
 * head's 16 weights and bias (ctr_unet_num_params() floats). */
typedef struct ctr_unet ctr_unet;
long long ctr_unet_num_params(void);
int ctr_unet_create(int device, ctr_unet** out);
int ctr_unet_destroy(ctr_unet* net);
/* Host parameters in / out (set also resets the Adam moments). */
int ctr_unet_set_params(ctr_unet* net, const float* host, long long n);
int ctr_unet_get_params(ctr_unet* net, float* host, long long n);
int ctr_unet_set_scales(ctr_unet* net, float s_x, float s_d, float s_out);
/* d_update = predicted update (device FP32 volume).  Asynchronous on stream. */
int ctr_unet_forward(ctr_unet* net, const float* d_x, const float* d_delta, float* d_update,
                     int Y, int X, int Z, void* stream);
/* Loss sum((net - d_target / s_out)^2) / (2 V) of the unscaled output and its gradient
 * with respect to every parameter (host_grads, may be NULL).  Synchronous. */
int ctr_unet_loss_grad(ctr_unet* net, const float* d_x, const float* d_delta,
                       const float* d_target, int Y, int X, int Z, float* host_grads,
                       double* loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on that loss. loss may be NULL. */
int ctr_unet_train_step(ctr_unet* net, const float* d_x, const float* d_delta,
                        const float* d_target, int Y, int X, int Z, float lr, double* loss);

/* ---- device memory helpers (so that hosts need no other GPU runtime) ---- */
int ctr_device_count(int* count);
int ctr_malloc(void** d_ptr, unsigned long long bytes, int device);
int ctr_free(void* d_ptr);
int ctr_memcpy_h2d(void* d_dst, const void* h_src, unsigned long long bytes);
int ctr_memcpy_d2h(void* h_dst, const void* d_src, unsigned long long bytes);
int ctr_memset(void* d_dst, int value, unsigned long long bytes);
int ctr_synchronize(int device);
/* Plain copy kernel of n floats (also the byte-counter calibration kernel). Asynchronous. */
int ctr_device_copy(float* d_dst, const float* d_src, unsigned long long n, void* stream);

/* n floats rounded to BF16 (nearest even) / n BF16 values widened to FP32. Asynchronous. */
int ctr_f32_to_bf16(unsigned short* d_dst, const float* d_src, unsigned long long n,
                    void* stream);
int ctr_bf16_to_f32(float* d_dst, const unsigned short* d_src, unsigned long long n,
                    void* stream);

/* d_out = d_a - d_b and d_y += a * d_x over n floats (the solvers' vector steps, for
 * hosts that keep iterates on the device). Asynchronous. */
int ctr_device_sub(float* d_out, const float* d_a, const float* d_b, unsigned long long n,
                   void* stream);
int ctr_device_axpy(float* d_y, const float* d_x, double a, unsigned long long n, void* stream);

/* Allocation-failure injection for tests (build addition): after n more successful
 * device allocations by the library, the next one fails with CTR_ERR_NOMEM (and the
 * injection turns itself off); n < 0 turns it off. */
int ctr_debug_fail_alloc_after(long long n);

/* Message for the last failing call on this thread ("" if none). */
const char* ctr_last_error(void);
/* Library version string. */
const char* ctr_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CTRECON_H */
