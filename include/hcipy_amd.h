/* C ABI of the MI355X (gfx950) Fourier-propagation library.
 *
 * Plain pointers and sizes only.  Device pointers are HIP device memory (for example the
 * data_ptr() of a torch tensor on a ROCm device); `stream` is a hipStream_t passed as void*.
 * Every function returns 0 on success or a nonzero status; hcipy_get_last_error() then returns a
 * thread-local message.  No function throws across the ABI.
 *
 * Complex arrays are interleaved (re, im) in the layout of numpy / torch complex64 and
 * complex128 arrays; a 2-D field of grid.shape (Ny, Nx) is row-major with x fastest
 * (hcipy/field/coordinates.py:671-675), a batch of fields is contiguous.
 */
#ifndef HCIPY_AMD_H_
#define HCIPY_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define HCIPY_OK 0
#define HCIPY_ERROR_INVALID_VALUE 1
#define HCIPY_ERROR_HIP 2
#define HCIPY_ERROR_OUT_OF_RANGE 3
#define HCIPY_ERROR_UNSUPPORTED 4

/* Arithmetic of the MFT kernels.  The dtype contract does not depend on it:
 * FP64 reads/writes complex128, FP32, F16X2 and BF16X3 read/write complex64. */
#define HCIPY_PREC_FP64 0   /* FP64 MFMA (complex128 fields)                                  */
#define HCIPY_PREC_FP32 1   /* FP32 MFMA, single-level FP32 accumulation (complex64 fields)   */
#define HCIPY_PREC_F16X2 2  /* hi/lo FP16 split MFMA, FP64 flush of partial sums (complex64)  */
#define HCIPY_PREC_BF16X3 3 /* three-way BF16 split MFMA (six products per k), FP64 flush of
                               the leading products (complex64)                                */

#define HCIPY_FORWARD 0
#define HCIPY_BACKWARD 1

/* Version of this ABI; bumped on any incompatible change. */
int hcipy_abi_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* hcipy_get_last_error(void);

typedef struct hcipy_mft_plan_s* hcipy_mft_plan;

/* Creates a matrix Fourier transform between separated Cartesian grids.
 * Replaces MatrixFourierTransform.__init__ + _compute_matrices
 * (hcipy/fourier/matrix_fourier_transform.py:42-65, 67-111).
 *
 * ndim      : 1 or 2.
 * x_in[nx]  : input x coordinates (FP64, host memory); y_in[ny] the y coordinates (2-D only).
 * u_out[nu] : output x coordinates; v_out[nv] the output y coordinates (2-D only).
 *   For a 1-D transform ny and nv must be 1 and y_in / v_out may be NULL.
 * precision : HCIPY_PREC_*.
 * The Fourier matrices exp(-i v y) and exp(-i x u) (and their adjoints) are generated on the
 * current device from FP64 phase products and stored in the precision's operand format. */
int hcipy_mft_plan_create(hcipy_mft_plan* plan, int ndim, int precision, const double* x_in, int64_t nx,
                          const double* y_in, int64_t ny, const double* u_out, int64_t nu, const double* v_out,
                          int64_t nv, void* stream);

/* Releases the device matrices of a plan.  Safe on NULL. */
int hcipy_mft_plan_destroy(hcipy_mft_plan plan);

/* Bytes of device workspace hcipy_mft_execute needs for `batch` fields (either direction). */
int hcipy_mft_workspace_size(hcipy_mft_plan plan, int64_t batch, size_t* bytes);

/* Bytes of device memory held by the plan's matrices. */
int hcipy_mft_plan_bytes(hcipy_mft_plan plan, size_t* bytes);

/* Workgroup tile shape of one contraction stage (stage 0 contracts the field's x axis, stage 1
 * its y axis).  No counterpart in the reference (its BLAS picks blocking internally); used by
 * hcipy_mft_autotune and by callers that pin a measured choice.  All shapes give identical
 * results up to summation order.  Plans start with HCIPY_TILES_AUTO (a size heuristic). */
#define HCIPY_TILES_AUTO -1
#define HCIPY_TILES_64X64_W2X2 0 /* 64x64 tile, 2x2 waves of 32x32                          */
#define HCIPY_TILES_64X32_W2X1 1 /* 64x32 tile, 2x1 waves, k-vectors split over 2 waves      */
#define HCIPY_TILES_64X64_K4 2   /* 64x64 tile per wave, k-vectors split over 4 waves (not FP64) */
#define HCIPY_TILES_64X32_K4 3   /* 64x32 tile per wave, k-vectors split over 4 waves (not FP64) */
#define HCIPY_TILES_128X64 4     /* 128x64 tile, 2x2 waves of 64x32, half-depth k-tiles (not FP64) */
#define HCIPY_TILES_64X128 5     /* 64x128 tile, 2x2 waves of 32x64, half-depth k-tiles (not FP64) */
#define HCIPY_TILES_64X64_W8 6   /* 64x64 tile, 8 waves: 2x2 waves of 32x32, k-vectors split over 2 */
#define HCIPY_TILES_64X32_W8 7   /* 64x32 tile, 8 waves: 2x1 waves of 32x32, k-vectors split over 4 (not FP64) */
/* F16X2 only: the constant matrix stays compact (complex) in LDS and each wave forms the Re and Im
 * operands in registers; tiles of 32 complex constant rows (= 64 real rows). */
#define HCIPY_TILES_C64X64_W4 8  /* 32x64 compact tile, 4 waves, k-vectors split over 2 */
#define HCIPY_TILES_C64X64_W8 9  /* 32x64 compact tile, 8 waves, k-vectors split over 4 */
#define HCIPY_TILES_C64X32_W4 10 /* 32x32 compact tile, 4 waves, k-vectors split over 4 */
#define HCIPY_TILES_C128X32_W8 11 /* 64x32 compact tile, 8 waves, k-vectors split over 4 */
/* F16X2 only: compact constant with the operand tiles loaded straight into LDS (three stages);
 * for contraction lengths that are not whole 128-element k-tiles the compact kernel runs. */
#define HCIPY_TILES_D64X64_W4 12 /* 32x64 compact tile, 4 waves */
#define HCIPY_TILES_D64X64_W8 13 /* 32x64 compact tile, 8 waves */
#define HCIPY_TILES_D64X32_W4 14 /* 32x32 compact tile, 4 waves */
#define HCIPY_TILES_D64X32_W8 15 /* 32x32 compact tile, 8 waves */
/* F16X2 parity form only (hcipy_mft_plan_set_parity), 64-deep k-tiles, two workgroups per CU: */
#define HCIPY_TILES_R64X64_W4 16 /* 64x64 tile, 2x2 waves of 32x32 */
#define HCIPY_TILES_R64X32_W4 17 /* 64x32 tile, 2x1 waves, k-vectors split over 2 */
#define HCIPY_TILES_R32X64_W4 18 /* 32x64 tile, 1x2 waves, k-vectors split over 2 */
/* Parity form, stage 2 only (F16X2; BF16X3 the second): 32 output rows x 32 (u', re / im)
 * columns of all four parity items per workgroup, one wave per item, and the recombination into
 * the output quadrants in the epilogue (no FP64 products in the workspace, no recombination
 * launch). */
#define HCIPY_TILES_RC32X32_W4 19 /* 64-deep k-tiles */
#define HCIPY_TILES_RC32X32_H 20  /* 32-deep k-tiles */
#define HCIPY_TILES_RC32X32_W8 21 /* 64-deep k-tiles (32-deep for split-BF16), 8 waves: two per item, each on half of every k-tile */
/* F16X2 parity form, either stage: 64x64 tile, 2x2 waves of 32x32, 32-deep k-tiles (41 KB of
 * LDS; a candidate for short contractions, e.g. in hcipy_mft_plan_set_tiles after timing) */
#define HCIPY_TILES_R64X64_H 22
/* F16X2 complex form, poke sets only (hcipy_poke_set_autotune, hcipy_poke_set_set_tiles): 128x64
 * tile as HCIPY_TILES_128X64 with 32-deep k-tiles, two workgroups per CU, for very many short
 * contractions */
#define HCIPY_TILES_128X64_H 23
int hcipy_mft_plan_set_tiles(hcipy_mft_plan plan, int direction, int stage, int tile_cfg);
int hcipy_mft_plan_get_tiles(hcipy_mft_plan plan, int direction, int stage, int* tile_cfg);

/* Parity form of a 2-D split-precision plan (HCIPY_PREC_F16X2 / BF16X3) whose four coordinate
 * axes are each symmetric about 0 (every sample x paired with a sample at -x, to 1e-12 of
 * max |x|, with at most one unpaired sample besides the origin: make_pupil_grid and
 * make_focal_grid grids of any size without shifts).  The field is split into its four x/y
 * parity components and each stage multiplies them by the real half-size cosine and sine
 * matrices: the same sum as the complex form (matrix_fourier_transform.py:171-172, 219-220) with
 * about a quarter of the multiply-adds; a final pass recombines the four quadrants.  Such plans
 * start with the parity form on; single executions
 * (hcipy_mft_execute / _ex / _noalloc, hcipy_mft_profile, hcipy_mft_autotune) use it, grouped
 * executions and poke sets keep the complex form.  set_parity(plan, 0) selects the complex form
 * (tile choices reset to HCIPY_TILES_AUTO on a switch; the parity form takes the shapes
 * 64X64_W2X2, 64X32_W2X1, 64X64_W8, 64X32_W8, 128X64 and, F16X2 only, 64X128 and the R* shapes); enabling it on an
 * ineligible plan is an error.  No counterpart in the reference. */
int hcipy_mft_plan_set_parity(hcipy_mft_plan plan, int enable);
int hcipy_mft_plan_get_parity(hcipy_mft_plan plan, int* available, int* enabled);

/* Fused parity execution: a single execution (hcipy_mft_execute[_ex], _noalloc) of an F16X2 plan
 * in the parity form whose stage shapes are R64X64_W4 and RC32X32_W4 (the defaults at
 * 1024^2 -> 512^2) runs as ONE persistent launch: the parity split, stage 1 and the recombining
 * stage 2 are tasks of one kernel that wait for each other through device counters held by the
 * plan, instead of three launches.  The results equal the three-launch path's (the same
 * arithmetic per tile).  On by default where available; set_fused(plan, 0) keeps three launches.
 * Because the counters belong to the plan, executions of one plan must not overlap in time
 * (executions on one stream never do; two streams running one plan at once need two plans).
 * get_fused reports availability (an F16X2 parity plan), the setting, and the number of waits
 * that gave up since the last call (0 unless a launch could not make progress; such a launch's
 * output is not valid).  get_fused synchronises the device when `timeouts` is not NULL.
 * No counterpart in the reference; it replaces the launch sequence of
 * hcipy/fourier/matrix_fourier_transform.py:171-172 (two zgemm/cgemm calls). */
int hcipy_mft_plan_set_fused(hcipy_mft_plan plan, int enable);
int hcipy_mft_plan_get_fused(hcipy_mft_plan plan, int* available, int* enabled, unsigned* timeouts);

/* Autotuner: for each stage of `direction`, times every tile shape available for the plan's
 * precision on the caller's buffers (`repeats` runs each, HIP events around the stage kernel),
 * pins the fastest on the plan and writes it to chosen[stage] (chosen may be NULL; chosen[1] is
 * HCIPY_TILES_AUTO for 1-D plans).  `out` and `workspace` are overwritten.  No counterpart in
 * the reference; it replaces the reference's measured planner choice
 * (make_fourier_transform(planner='measure'), hcipy/fourier/fourier_transform.py:217-300) at the
 * level of kernel tiles. */
int hcipy_mft_autotune(hcipy_mft_plan plan, int direction, const void* in, void* out, int64_t batch, int output,
                       void* workspace, size_t workspace_bytes, int repeats, int* chosen, void* stream);

/* Runs `batch` transforms on `stream`.
 * Replaces MatrixFourierTransform.forward / .backward (matrix_fourier_transform.py:132-176,
 * 178-224) for scalar grid weights:
 *   forward : out[b] = alpha * M1 . in[b] . M2          in: (ny, nx) -> out: (nv, nu)
 *   backward: out[b] = alpha * M1^H . in[b] . M2^H      in: (nv, nu) -> out: (ny, nx)
 * The caller folds the grid weight (w_in, or w_out / (2 pi)^ndim for backward) and any
 * propagator normalisation into alpha.  `in` and `out` must not overlap.  `workspace` must hold
 * hcipy_mft_workspace_size(plan, batch) bytes and is not read before being written.  Calls that
 * share a workspace must be ordered on one stream.  No allocation happens in this call. */
int hcipy_mft_execute(hcipy_mft_plan plan, int direction, const void* in, void* out, int64_t batch,
                      double alpha_re, double alpha_im, void* workspace, size_t workspace_bytes, void* stream);

/* Grouped execution: many transforms, with different plans (matrices, sizes, e.g. one per
 * wavelength of a broadband source or per pupil size) and batches, as ONE launch per contraction
 * stage (plus the per-item data split of split-FP16 plans).  Replaces the reference's loop over
 * wavelengths / instances of MatrixFourierTransform.forward and .backward
 * (hcipy/fourier/matrix_fourier_transform.py:132-224 called per AgnosticOpticalElement instance,
 * hcipy/optics/optical_element.py:297-380); the stage kernels are the same as hcipy_mft_execute's
 * (one tile shape per launch, chosen from the items' longest contraction and total size).  All
 * items share one precision and
 * live on the group's device; every item has its own workspace (hcipy_mft_workspace_size).
 * A group owns a device table of stage arguments and a pinned host mirror (no allocation in
 * hcipy_mft_execute_grouped).  Calls on one group are ordered on one stream.  Items that read the
 * same input with the same rows and contraction length share one data split (split-FP16).
 * A call recorded under stream capture copies its table from a pinned snapshot of its own, so every
 * replay sees the table of its capture.  A group holds 64 snapshots (fewer for very large
 * capacities), taken in turn by captured calls; once all are taken a captured call fails until
 * hcipy_mft_group_release_captures returns them.  Graphs that recorded calls of a group must not be
 * replayed after its snapshots are released or the group is destroyed (they read its tables). */
typedef struct hcipy_mft_group_s* hcipy_mft_group;
typedef struct {
  hcipy_mft_plan plan;
  int direction;             /* HCIPY_FORWARD / HCIPY_BACKWARD */
  const void* in;
  void* out;
  int64_t batch;
  double alpha_re, alpha_im;
  int output;                /* HCIPY_OUTPUT_* */
  void* workspace;
  size_t workspace_bytes;
} hcipy_mft_group_item;
int hcipy_mft_group_create(hcipy_mft_group* group, int capacity);
int hcipy_mft_group_destroy(hcipy_mft_group group);
int hcipy_mft_execute_grouped(hcipy_mft_group group, const hcipy_mft_group_item* items, int n_items, void* stream);
/* Hands every capture snapshot of the group back (held, may be NULL: how many were taken).  The
 * caller states that no graph recorded before this call is replayed again (destroy or re-capture
 * them).  Host-only; no counterpart in the reference, whose loop over instances is not captured. */
int hcipy_mft_group_release_captures(hcipy_mft_group group, int* held);

/* Output kinds of hcipy_mft_execute_ex. */
#define HCIPY_OUTPUT_FIELD 0         /* complex field in the plan's dtype (complex128 / complex64) */
#define HCIPY_OUTPUT_FIELD_C128 1    /* complex128 field from any precision                       */
#define HCIPY_OUTPUT_INTENSITY 2     /* |out|^2 from the FP64 epilogue value: float64 for FP64
                                        plans, float32 otherwise                                  */
#define HCIPY_OUTPUT_INTENSITY_F64 3 /* |out|^2 as float64                                        */

/* hcipy_mft_execute with a choice of output: the field, or the intensity |out|^2 evaluated in the
 * epilogue before any rounding of the field to the output dtype.  The intensity output replaces
 * FraunhoferPropagator.forward followed by Wavefront.intensity (hcipy/propagation/fraunhofer.py:56-71,
 * hcipy/optics/wavefront.py:113-119) in one pass. */
int hcipy_mft_execute_ex(hcipy_mft_plan plan, int direction, const void* in, void* out, int64_t batch,
                         double alpha_re, double alpha_im, int output, void* workspace, size_t workspace_bytes,
                         void* stream);

/* Bytes of the static device scratch behind hcipy_mft_execute_noalloc (32 MiB per device). */
#define HCIPY_NOALLOC_SCRATCH_BYTES (32 * 1024 * 1024)

/* hcipy_mft_execute_ex without a workspace argument and without any allocation: the split data
 * and the intermediate go to a static device array of HCIPY_NOALLOC_SCRATCH_BYTES that every
 * device loading the library holds until the process ends.  No counterpart in the reference; it
 * is for callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Fails with HCIPY_ERROR_OUT_OF_RANGE, launching nothing, when
 * the workspace this direction needs (at most hcipy_mft_workspace_size(plan, batch)) exceeds the
 * scratch (e.g. 1024^2 -> 512^2 split-FP16 forward fits batch 2).  All calls on one device share its scratch: calls that overlap on one device race, so
 * order them on one stream.  The first call on a device looks the scratch up (no allocation). */
int hcipy_mft_execute_noalloc(hcipy_mft_plan plan, int direction, const void* in, void* out, int64_t batch,
                              double alpha_re, double alpha_im, int output, void* stream);

/* Runs hcipy_mft_execute_ex once on `stream`, then launches each of its kernels alone `repeats`
 * times back to back between two HIP events, waits, and stores the mean duration per launch in
 * milliseconds of: kernel_ms[0] the data split (split-FP16 only, else ~0), kernel_ms[1] stage 1,
 * kernel_ms[2] stage 2 (all batch slices).  `out` and `workspace` are overwritten.
 * Measurement aid with no counterpart in the reference (whose MFT is timed by _time_it,
 * hcipy/fourier/fourier_transform.py:185-215); not for production loops. */
int hcipy_mft_profile(hcipy_mft_plan plan, int direction, const void* in, void* out, int64_t batch, double alpha_re,
                      double alpha_im, int output, void* workspace, size_t workspace_bytes, int repeats,
                      float* kernel_ms, void* stream);
/* As hcipy_mft_profile, but each step is timed inside whole executions, with the cache state the
 * step before it leaves: `repeats` back-to-back executions of the split alone, of split and stage
 * 1, and of all steps are each timed between two HIP events, and kernel_ms[k] is the difference of
 * consecutive totals per execution (no events between the launches of one execution).  With the
 * recombining stage-2 shapes of the parity form kernel_ms[2] covers that one launch, otherwise
 * stage 2 and the recombination.  Measurement aid with no counterpart in the reference. */
int hcipy_mft_profile_pipeline(hcipy_mft_plan plan, int direction, const void* in, void* out, int64_t batch,
                               double alpha_re, double alpha_im, int output, void* workspace,
                               size_t workspace_bytes, int repeats, float* kernel_ms, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Pointwise kernels of the propagation chain.  `batch` fields of `n` samples each lie
 * contiguously; a modulating array is broadcast over the batch when its stride is 0, or read
 * per field with stride n.  Complex64 fields are combined in FP32, complex128 in FP64. */

#define HCIPY_DTYPE_C64 0
#define HCIPY_DTYPE_C128 1
#define HCIPY_DTYPE_F32 2
#define HCIPY_DTYPE_F64 3

/* out = (x - y) * a    (y and a optional; conj(a) if conjugate_a; x NULL reads as zero, for
 * the adjoint -g conj(a) with respect to y).
 * Replaces Apodizer.forward/backward (hcipy/optics/apodization.py:25-35) with y = NULL, the
 * Babinet step F -= M F of the Lyot coronagraph (hcipy/coronagraphy/lyot.py:80) with a = 1-M,
 * and E - P^dagger F followed by the Lyot stop (lyot.py:87-90) in one pass.  out may alias x. */
int hcipy_pw_sub_mul(const void* x, const void* y, const void* a, int a_dtype, int64_t a_stride, int conjugate_a,
                     void* out, int field_dtype, int64_t n, int64_t batch, void* stream);

/* out = x * exp(i * scale * phi) with the phase evaluated in FP64.
 * Replaces PhaseApodizer (hcipy/optics/apodization.py:37-62, scale = +-1) and the
 * DeformableMirror phase exp(+-2ik s) (hcipy/optics/deformable_mirror.py:200-238,
 * scale = +-2k, phi = surface).  out may alias x. */
int hcipy_pw_phase(const void* x, const void* phi, int phi_dtype, int64_t phi_stride, double scale, void* out,
                   int field_dtype, int64_t n, int64_t batch, void* stream);
/* The same with field b scaled by scale * row_scale[b] (row_scale: `batch` doubles in device
 * memory, or NULL): one launch for the wavelengths of a broadband source, each with its own
 * 2k = 4 pi / lambda.  No counterpart in the reference (it evaluates one wavelength per call). */
int hcipy_pw_phase_rows(const void* x, const void* phi, int phi_dtype, int64_t phi_stride, const double* row_scale,
                        double scale, void* out, int field_dtype, int64_t n, int64_t batch, void* stream);

/* out[i] = |x[i]|^2 (float32 for complex64, float64 for complex128).
 * Replaces Wavefront.I for scalar fields (hcipy/optics/wavefront.py:113-119). */
int hcipy_pw_intensity(const void* x, int field_dtype, void* out, int64_t total, void* stream);

/* Bytes of workspace hcipy_pw_power needs for `batch` fields. */
int hcipy_pw_power_workspace(int64_t batch, size_t* bytes);

/* out[b] = weight * sum_i mask[i] |x_b[i]|^2, accumulated in FP64 (mask optional).
 * Replaces Wavefront.total_power (hcipy/optics/wavefront.py:332-360, weight = grid weight)
 * and the dark-zone sums behind get_mean_intensity_in_roi (hcipy/metrics/contrast.py:37-52).
 * Deterministic two-pass reduction; `out` is device memory. */
int hcipy_pw_power(const void* x, int field_dtype, const void* mask, int mask_dtype, int64_t mask_stride,
                   double weight, int64_t n, int64_t batch, double* out, void* workspace, size_t workspace_bytes,
                   void* stream);

/* out[b] = mean of x[b][i] over the i with mask[b * mask_stride + i] != 0 (all i if mask is
 * NULL), for real float32 / float64 values (HCIPY_DTYPE_F32 / F64), summed in FP64 in a fixed
 * order; NaN for an empty region.  Replaces get_mean_intensity_in_roi
 * (hcipy/metrics/contrast.py:37-52, np.mean(img[mask])).  `out` is device memory; the workspace
 * holds hcipy_pw_masked_mean_workspace(batch) bytes. */
int hcipy_pw_masked_mean_workspace(int64_t batch, size_t* bytes);
int hcipy_pw_masked_mean(const void* x, int dtype, const void* mask, int mask_dtype, int64_t mask_stride, int64_t n,
                         int64_t batch, double* out, double* count, void* workspace, size_t workspace_bytes,
                         void* stream);
/* (count, device, may be NULL: the number of selected samples per batch item.) */

/* ---------------------------------------------------------------------------------------------
 * Adjoints (vector-Jacobian products) of the pointwise operations, with PyTorch's convention
 * for complex gradients (g = dL/d conj(output)).  They serve the torch.autograd.Function wrappers
 * of the package; the reference has no differentiation (its backward() of an optical element is
 * the adjoint propagation these reproduce for the linear operations). */

/* out[b][i] = g[b] / count[b] where mask != 0, else 0: adjoint of hcipy_pw_masked_mean. */
int hcipy_pw_masked_mean_adjoint(const void* mask, int mask_dtype, int64_t mask_stride, const double* g,
                                 const double* count, void* out, int out_dtype, int64_t n, int64_t batch,
                                 void* stream);

/* Gradient with respect to phi of y = x exp(i scale phi) (hcipy_pw_phase) given y and g:
 * out = -scale Im(conj(g) y) in FP64, summed over the batch when phi_stride is 0. */
int hcipy_pw_phase_adjoint(const void* y, const void* g, int field_dtype, double scale, double* out,
                           int64_t phi_stride, int64_t n, int64_t batch, void* stream);
/* Adjoint of hcipy_pw_phase_rows: out = -scale sum_b row_scale[b] Im(conj(g_b) y_b) for a shared
 * phase (phi_stride 0), -scale row_scale[b] Im(conj(g_b) y_b) per field otherwise. */
int hcipy_pw_phase_rows_adjoint(const void* y, const void* g, int field_dtype, const double* row_scale, double scale,
                                double* out, int64_t phi_stride, int64_t n, int64_t batch, void* stream);

/* Gradient with respect to x of I = |x|^2 (hcipy_pw_intensity): out = 2 g x, g real. */
int hcipy_pw_intensity_adjoint(const void* x, int field_dtype, const void* g, int g_dtype, int64_t g_stride, void* out,
                               int64_t n, int64_t batch, void* stream);

/* Per-pixel matrix product of a (p x q) matrix field J and a field E of q x K components on n
 * pixels: out[i][k][px] = sum_j J[i][j][px] E[j][k][px], layouts [row][col][pixel].  With
 * adjoint != 0 the conjugate transpose J^H is applied (E then has p x K components, out q x K).
 * Replaces field_dot of a Jones matrix with a Jones-vector or Jones-matrix field
 * (hcipy/field/field.py field_dot, used by JonesMatrixOpticalElement.forward/backward,
 * hcipy/optics/polarization.py:44-116, and VectorVortexCoronagraph.forward,
 * hcipy/coronagraphy/vortex.py:162-199).  J may be complex64 or complex128. */
int hcipy_pw_field_dot(const void* J, int j_dtype, int p, int q, int adjoint, const void* E, int field_dtype, int64_t K,
                       void* out, int64_t n, void* stream);

/* out = x - sum_j ys[j] (j = 0 .. k-1, k <= 16, subtracted in that order) over n complex values;
 * ys is a host array of device pointers and out may equal x.  Adds the level terms of
 * MultiScaleCoronagraph.forward/backward (hcipy/coronagraphy/multi_scale.py:108-182) onto the
 * Lyot-plane field in one pass (the level transforms are computed with a negated scale). */
int hcipy_pw_sub_many(const void* x, const void* const* ys, int k, void* out, int field_dtype, int64_t n,
                      void* stream);

/* out = x exp(i (ax u + ay v)) for `batch` fields on a regular nx x ny grid with coordinates
 * u = zero_x + ix delta_x, v = zero_y + iy delta_y (x fastest), phase in FP64; out may equal x.
 * Replaces the shift of an FFT spectral-noise realization, C *= exp(-1j * (s_x u + s_y v))
 * (hcipy/util/spectral_noise.py:177-187; ax = -s_x, ay = -s_y).  A 1-D grid has ny = 1, ay = 0. */
int hcipy_pw_phase_ramp(const void* x, int field_dtype, int nx, int ny, double zero_x, double delta_x, double zero_y,
                        double delta_y, double ax, double ay, void* out, int64_t batch, void* stream);

/* Stokes parameters (I, Q, U, V) per pixel of a polarized wavefront: Wavefront.I/.Q/.U/.V
 * (hcipy/optics/wavefront.py:112-235).  With vector_field != 0, E is a Jones-vector field
 * (layout [2][n]) and stokes is ignored; otherwise E is the 2 x 2 Jones-matrix field (layout
 * [2][2][n]) of a partially polarized wavefront acting on the input Stokes vector
 * stokes[0..3].  component 0..3 writes that parameter to out[n]; -1 writes all four to
 * out[4][n] (Wavefront.stokes_vector).  FP64 arithmetic; float32 out for complex64 fields and
 * float64 for complex128. */
int hcipy_pw_stokes(const void* E, int field_dtype, int vector_field, const double* stokes, int component, void* out,
                    int64_t n, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Deformable mirror with Gaussian influence functions on a regular pupil grid, evaluated
 * without forming the influence matrix.  Replaces make_gaussian_influence_functions +
 * DeformableMirror.surface (hcipy/optics/deformable_mirror.py:43-93, 255-267;
 * hcipy/mode_basis/gaussian_pokes.py:5-54; ModeBasis.linear_combination,
 * hcipy/mode_basis/mode_basis.py:239-258).  Influence function of actuator a at pixel p:
 *   q = R(-z_tilt) . diag(1/cos(y_tilt), 1/cos(x_tilt)) . (p + dither),  r2 = |q - m_a|^2,
 *   g = exp(-r2 / (2 sigma^2)) - exp(-radius^2 / (2 sigma^2)) for r2 <= radius^2, else 0,
 *   IF_a(p) = mean of g over oversampling^2 dithers / (cos(x_tilt) cos(y_tilt)),
 * actuators m_a on an n_act x n_act grid (x fastest) of the given origin and pitch.
 * The reference's defaults: sigma = pitch / sqrt(-2 ln crosstalk), radius = 3 pitch. */
typedef struct {
  int64_t nx, ny;          /* pupil grid samples, x fastest */
  double x0, y0, dx, dy;   /* pupil grid coordinates x_i = x0 + i dx, y_j = y0 + j dy */
  int64_t n_act;           /* actuators across (n_act^2 actuators) */
  double act_x0, act_y0;   /* position of actuator (0, 0) */
  double pitch;            /* actuator spacing */
  double sigma, radius;    /* Gaussian width and support radius */
  double x_tilt, y_tilt, z_tilt;
  int64_t oversampling;    /* 1..16 */
} hcipy_dm_gaussian_desc;

/* surface[b][pixel] = sum_a actuators[b][a] IF_a(pixel), FP64, for `batch` actuator vectors. */
int hcipy_dm_gaussian_surface(const hcipy_dm_gaussian_desc* desc, const double* actuators, double* surface,
                              int64_t batch, void* stream);

/* grad_actuators[b][a] = sum_pixel grad_surface[b][pixel] IF_a(pixel): the adjoint of the
 * surface map, i.e. the transposed influence matrix (the reference applies
 * influence_functions.transformation_matrix.T for DM Jacobians).  Deterministic order. */
int hcipy_dm_gaussian_surface_adjoint(const hcipy_dm_gaussian_desc* desc, const double* grad_surface,
                                      double* grad_actuators, int64_t batch, void* stream);

/* Poke set: focal-plane intensities of every single-actuator poke of a deformable mirror, for
 * several propagation items (wavelengths with their own Fourier matrices, pupils of different
 * sizes), as one set of grouped launches.  Replaces the reference's DM-Jacobian loop: poke one
 * actuator (DeformableMirror.actuators, hcipy/optics/deformable_mirror.py:200-263, Gaussian pokes
 * of hcipy/mode_basis/gaussian_pokes.py:5-54), propagate (FraunhoferPropagator.forward through the
 * per-wavelength instance cache, hcipy/optics/optical_element.py:297-380), take the intensity.
 * Item k with plan P (forward), base pupil field E and scale alpha gives
 *   base_out            = |alpha P(E)|^2
 *   poke_out[a]         = |alpha P(E exp(i phase_scale IF_a))|^2      for every actuator a,
 * computed as |alpha P(E) + alpha P_w(dE_a)|^2 where dE_a = E (exp(i phase_scale IF_a) - 1) vanishes
 * outside the pixel window of IF_a's support and P_w is the transform restricted to that window
 * (the same Fourier matrices, offset to the window's rows and columns): linearity makes this the
 * same sum with 8x less arithmetic at 1024^2 with 32 x 32 actuators.  Intensities are float32
 * (float64 for FP64 plans), [actuator][nv][nu].
 * All plans share one precision, are 2-D and live on the current device; the DM descriptor of
 * an item describes its plan's input grid.  The set owns all its device memory (workspaces,
 * base focal fields, launch tables), allocated at creation: hcipy_poke_set_execute neither
 * allocates nor copies, reads the item fields as they are at that point of the stream and can be
 * captured in a graph.  Creation refuses a capturing stream. */
typedef struct hcipy_poke_set_s* hcipy_poke_set;
typedef struct {
  hcipy_mft_plan plan;         /* forward transform pupil -> focal (uv) grid                 */
  const void* field;           /* base pupil field, complex64 (complex128 for FP64 plans)     */
  double alpha_re, alpha_im;   /* transform scale (grid weight x propagator normalisation)   */
  hcipy_dm_gaussian_desc dm;   /* influence functions on the plan's input grid               */
  double phase_scale;          /* phase of the poke on a unit influence function: 2 k h      */
  void* base_out;              /* base intensity (nv * nu) or NULL                           */
  void* poke_out;              /* n_act^2 outputs [actuator][nv * nu]                        */
  int output;                  /* HCIPY_POKE_INTENSITY or HCIPY_POKE_FIELD_CHANGE            */
} hcipy_poke_item;
/* Outputs of a poke: the poked focal intensity (float32, float64 for FP64), or the change of the
 * focal field alpha P_w(dE_a) (complex64, complex128 for FP64) -- the columns of the DM Jacobian
 * times the poke height, as an electric-field-conjugation loop uses them. */
#define HCIPY_POKE_INTENSITY 0
#define HCIPY_POKE_FIELD_CHANGE 1
int hcipy_poke_set_create(hcipy_poke_set* set, const hcipy_poke_item* items, int n_items, void* stream);
int hcipy_poke_set_bytes(hcipy_poke_set set, size_t* bytes);
/* Pokes of the set and the real flops (8 per complex multiply-add) of one execution as run (base
 * transforms plus windowed pokes), and of the same results by full transforms of every poked field. */
int hcipy_poke_set_info(hcipy_poke_set set, int64_t* n_pokes, double* flops_windowed, double* flops_full);
int hcipy_poke_set_execute(hcipy_poke_set set, void* stream);
/* Autotuner of a poke set: times `repeats` executions with every tile shape available to grouped
 * launches, first for the stage-1 launch of the pokes, then for stage 2, keeps the fastest and
 * writes them to chosen[0..1] (may be NULL).  Runs executions (the outputs are overwritten);
 * refuses a capturing stream.  No counterpart in the reference. */
int hcipy_poke_set_autotune(hcipy_poke_set set, int repeats, int* chosen, void* stream);
/* Tile shapes of the stage-1 and stage-2 launches of the pokes (HCIPY_TILES_*, or HCIPY_TILES_AUTO
 * for the default), any shape hcipy_poke_set_autotune may choose; HCIPY_ERROR_INVALID_VALUE for
 * others.  Uploads the set's launch tables with synchronous copies: not under stream capture.  No
 * counterpart in the reference. */
int hcipy_poke_set_set_tiles(hcipy_poke_set set, int stage1_tiles, int stage2_tiles);
int hcipy_poke_set_destroy(hcipy_poke_set set);

/* Tikhonov-regularised inverse of a real m x n matrix M (row-major, device, FP64):
 *   out (n x m, row-major) = V diag(s / (s^2 + (rcond s_max)^2)) U^T = (M^T M + lam^2 I)^-1 M^T,
 * lam = rcond * s_max, s_max the largest singular value (written to *s_max, host, may be NULL).
 * Replaces inverse_tikhonov (hcipy/util/matrix_inversion.py:71-100), which takes an SVD: here
 * M^T M runs on the FP64 MFMA stage kernel, s_max^2 comes from power iteration on it, and the
 * regularised normal equations are solved by a blocked device Cholesky factorisation.  Relative
 * accuracy about 1e-16 / rcond^2 (the conditioning of the normal equations); rcond must keep
 * M^T M + lam^2 I positive definite in FP64.  Synchronises with the host (convergence checks).
 * The workspace holds hcipy_inverse_tikhonov_workspace(m, n) bytes. */
int hcipy_inverse_tikhonov_workspace(int64_t m, int64_t n, size_t* bytes);
int hcipy_inverse_tikhonov(const double* M, int64_t m, int64_t n, double rcond, double* out, double* s_max,
                           void* workspace, size_t workspace_bytes, void* stream);

/* Deformable mirror whose influence function is a sampled actuator shape: the Xinetics DM5V2
 * shape of make_xinetics_influence_functions (hcipy/optics/deformable_mirror.py:95-138).
 * IF_a(p) = bilinear interpolation of `table` at q - m_a (zero outside the table's extent), with
 * q = R(-z_tilt) . diag(1/cos(y_tilt), 1/cos(x_tilt)) . p, times cos(y_tilt) / cos(x_tilt) (the
 * reference's "/ cos(x_tilt) * cos(y_tilt)").  The table is a device array [table_ny][table_nx]
 * on the coordinates table_x0 + i table_dx, table_y0 + j table_dy, at most 128 x 128 samples. */
typedef struct {
  int64_t nx, ny;
  double x0, y0, dx, dy;
  int64_t n_act;
  double act_x0, act_y0;
  double pitch;
  double x_tilt, y_tilt, z_tilt;
  int64_t table_nx, table_ny;
  double table_x0, table_y0, table_dx, table_dy;
  const double* table;
} hcipy_dm_table_desc;

/* surface[b][pixel] = sum_a actuators[b][a] IF_a(pixel), FP64. */
int hcipy_dm_table_surface(const hcipy_dm_table_desc* desc, const double* actuators, double* surface, int64_t batch,
                           void* stream);
/* grad_actuators[b][a] = sum_pixel grad_surface[b][pixel] IF_a(pixel) (transposed influence matrix). */
int hcipy_dm_table_surface_adjoint(const hcipy_dm_table_desc* desc, const double* grad_surface, double* grad_actuators,
                                   int64_t batch, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Fourier filter: zero-pad -> 2-D FFT -> transfer function -> inverse 2-D FFT -> crop.
 * Replaces FourierFilter (hcipy/fourier/fourier_operations.py:8-151) for scalar transfer
 * functions and power-of-two internal grids, the engine of AngularSpectrumPropagator
 * (hcipy/propagation/angular_spectrum.py:45-152) and FresnelPropagator
 * (hcipy/propagation/fresnel.py:43-164).  Field: ny x nx (x fastest) centred in the NY x NX
 * internal grid at int(N/2) - int(n/2) (fast_fourier_transform.py:262-266). */

typedef struct hcipy_filter_plan_s* hcipy_filter_plan;

/* nx, ny: field size; NX, NY: internal (padded) size, powers of two in [4, 8192].
 * dtype: HCIPY_DTYPE_C64 or HCIPY_DTYPE_C128 (fields and transfer function). */
int hcipy_filter_plan_create(hcipy_filter_plan* plan, int nx, int ny, int NX, int NY, int dtype, void* stream);

int hcipy_filter_plan_destroy(hcipy_filter_plan plan);

/* Sets the transfer function from an NY x NX device array sampled on the centred internal grid
 * (the reference ifftshifts it, fourier_operations.py:41-52; so does this call). */
int hcipy_filter_set_transfer_function(hcipy_filter_plan plan, const void* tf_centred, int tf_dtype, void* stream);

/* Generates a near-field transfer function on the device in FP64, averaged over
 * oversampling^2 dithered sub-grids (hcipy/field/util.py:606-698).  The internal grid has
 * separated coordinates i * delta + zero.  kind 0: angular spectrum
 * exp(i sqrt(k^2 - |k|^2 + 0j) z) (angular_spectrum.py:71-78); kind 1: Fresnel
 * exp(i k z) exp(-i z |k|^2 / (2k)) (fresnel.py:69-80). */
int hcipy_filter_nearfield_transfer_function(hcipy_filter_plan plan, int kind, double delta_x, double zero_x,
                                             double delta_y, double zero_y, double wavenumber, double distance,
                                             int oversampling, void* stream);

/* The same transfer function on an nx x ny frequency grid (separated coordinates i * delta + zero),
 * written as complex128 in grid (centred) order: for filters whose padded grid is not a power
 * of two, which transform with the zoom FFT or the MFT and multiply by it pointwise
 * (angular_spectrum.py:71-78, fresnel.py:69-80). */
int hcipy_nearfield_transfer_function(int kind, int nx, int ny, double delta_x, double zero_x, double delta_y,
                                      double zero_y, double wavenumber, double distance, int oversampling, void* out,
                                      void* stream);

/* Impulse response of a near-field propagator on an nx x ny spatial grid (separated coordinates
 * i * delta + zero), FP64, averaged over oversampling^2 dithered sub-grids, written as complex128
 * in grid order.  The impulse-response branch of the propagators Fourier transforms it into the
 * transfer function when the grid step is below lambda |z| / L (hcipy/propagation/angular_spectrum.py:55-69,
 * hcipy/propagation/fresnel.py:50-67).  kind 0: Rayleigh-Sommerfeld
 * (z / r) / (2 pi) exp(i k r) (1 / r^2 - i k / r); kind 1: Fresnel
 * exp(i k z) / (i lambda z) exp(i k (x^2 + y^2) / (2 z)). */
int hcipy_nearfield_impulse_response(int kind, int nx, int ny, double delta_x, double zero_x, double delta_y,
                                     double zero_y, double wavenumber, double wavelength, double distance,
                                     int oversampling, void* out, void* stream);

/* Bytes of workspace hcipy_filter_execute needs for `batch` fields. */
int hcipy_filter_workspace_size(hcipy_filter_plan plan, int64_t batch, size_t* bytes);

/* out[b] = crop(ifft2(fft2(pad(in[b])) * tf)) (conj(tf) when adjoint != 0), any batch size.
 * in and out are batch x ny x nx complex arrays of the plan's dtype and must not overlap. */
int hcipy_filter_execute(hcipy_filter_plan plan, int adjoint, const void* in, void* out, int64_t batch,
                         void* workspace, size_t workspace_bytes, void* stream);

/* k <= 8 filters (plans with the same grids, dtype and device, each with its transfer function)
 * of the same input: outs[t] = plan t's filter of `in`, sharing the padded row FFTs and the
 * forward column FFTs.  The element filters of a matrix transfer function applied to one field
 * (FourierFilter with a matrix transfer function, hcipy/fourier/fourier_operations.py:120-131,
 * on a field times the identity, as VectorVortexCoronagraph's level 0 sees an unpolarized
 * wavefront, hcipy/coronagraphy/vortex.py:178-185).  Outputs must not alias the input. */
int hcipy_filter_workspace_size_multi(const hcipy_filter_plan* plans, int k, int64_t batch, size_t* bytes);
int hcipy_filter_execute_multi(const hcipy_filter_plan* plans, int k, int adjoint, const void* in, void* const* outs,
                               int64_t batch, void* workspace, size_t workspace_bytes, void* stream);

/* Matrix transfer function (n_out x n_in <= 4 x 4 plans with the same grids, dtype and device,
 * plans[i * n_in + j] holding element (i, j)): out[i][b] = sum_j filter_ij(in[j][b]), with
 * conj(tf) when conj != 0 (the adjoint passes the transposed plan matrix).  in is
 * [n_in][batch][ny][nx], out [n_out][batch][ny][nx], batch <= 65535.  Each input is
 * transformed once and each output inverse-transformed once.  Replaces FourierFilter with a
 * matrix transfer function, field_dot per frequency (hcipy/fourier/fourier_operations.py:120-131). */
int hcipy_filter_workspace_size_matrix(hcipy_filter_plan plan, int n_out, int n_in, int64_t batch, size_t* bytes);
int hcipy_filter_execute_matrix(const hcipy_filter_plan* plans, int n_out, int n_in, int conj, const void* in, void* out,
                                int64_t batch, void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Zoom FFT: the Fourier transform between two regular Cartesian grids of any spacing, as one
 * Bluestein chirp-z transform per axis.  Replaces ZoomFastFourierTransform
 * (hcipy/fourier/zoom_fast_fourier_transform.py:10-219) and ChirpZTransform
 * (hcipy/fourier/chirp_z_transform.py:7-105).
 *
 * Grids are given per axis in (x, y) order: sizes, spacings and first coordinates.  Forward maps
 * a field on the input grid (layout [ny][nx]) to alpha * sum_j f_j exp(-i x_j . u_k) on the
 * output grid ([my][mx]); backward maps the output grid to the input grid with exp(+i u . x).
 * The weights are the caller's alpha (the reference multiplies by input_grid.weights forward
 * and output_grid.weights / (2 pi)^ndim backward).  Each axis needs n_in + n_out - 1 <= 8192.
 * dtype: HCIPY_DTYPE_C64 or HCIPY_DTYPE_C128 fields.  precision: HCIPY_PREC_FP64 (FP64
 * chirps, line FFTs and intermediate; for complex64 fields only the loads and stores are FP32)
 * or, for complex64 fields, HCIPY_PREC_FP32 (all FP32, chirps and filters rounded from FP64).
 * Set-up (chirps, filter FFTs) runs on `stream`. */
typedef struct hcipy_zoom_plan_s* hcipy_zoom_plan;
int hcipy_zoom_plan_create(hcipy_zoom_plan* plan, int ndim, const int* n_in, const double* delta_in,
                           const double* zero_in, const int* n_out, const double* delta_out, const double* zero_out,
                           int dtype, int precision, void* stream);
/* The same plan for an FFT pair of grids: per axis delta_in * delta_out = 2 pi / fft_n[d] (to
 * 1e-9, else an error), fft_n the FFT length of the internal grid (the output grid is that FFT
 * grid, cropped and shifted as make_fft_grid makes it).  The quadratic chirp phases then use the
 * exact 2 pi / fft_n, reduced in integers, so the result is the exact DFT of FastFourierTransform
 * with its shift ramps (fourier/fast_fourier_transform.py:280-406) and its rounding does not grow
 * with the chirp's t^2.  Execution, workspace and destruction as for the general plan. */
int hcipy_zoom_plan_create_fft(hcipy_zoom_plan* plan, int ndim, const int* n_in, const double* delta_in,
                               const double* zero_in, const int* n_out, const double* delta_out,
                               const double* zero_out, const int* fft_n, int dtype, int precision, void* stream);
int hcipy_zoom_plan_destroy(hcipy_zoom_plan plan);
/* Bytes of workspace (the intermediate of the two axis passes) for `batch` fields; 0 in 1-D. */
int hcipy_zoom_workspace_size(hcipy_zoom_plan plan, int64_t batch, size_t* bytes);
/* direction 0: forward (zoom_fast_fourier_transform.py:70-92), 1: backward (:94-116).  `batch`
 * fields of the plan's dtype, contiguous; in and out must not alias. */
int hcipy_zoom_execute(hcipy_zoom_plan plan, int direction, const void* in, void* out, int64_t batch, double alpha_re,
                       double alpha_im, void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Direct FFT between a grid and its FFT grid for a power-of-two internal length (per axis
 * fft_n[d] in [4, 8192]).  Replaces FastFourierTransform.forward / .backward
 * (hcipy/fourier/fast_fourier_transform.py:280-406): the reference zero-pads the field into an
 * fft_n buffer, fftshifts, runs numpy's FFT and multiplies shift ramps; this plan runs one
 * length-fft_n line FFT per row and per column of the unpadded field, with the input ramp
 * exp(-i u0 dx n) applied on load and the output ramps exp(-i u0 x0) exp(-i du x0 k) on store,
 * keeping only the fft_n-grid bins the output grid holds.  Arguments as for
 * hcipy_zoom_plan_create_fft (same grids, delta_in * delta_out = 2 pi / fft_n[d] to 1e-9); the
 * output grid's first coordinate may sit anywhere on the FFT lattice.  precision: HCIPY_PREC_FP64
 * (FP64 line FFTs and intermediate; complex64 fields load and store FP32) or, for complex64
 * fields, HCIPY_PREC_FP32.  Set-up runs on `stream` and is refused under stream capture. */
typedef struct hcipy_fft_plan_s* hcipy_fft_plan;
int hcipy_fft_plan_create(hcipy_fft_plan* plan, int ndim, const int* n_in, const double* delta_in,
                          const double* zero_in, const int* n_out, const double* delta_out, const double* zero_out,
                          const int* fft_n, int dtype, int precision, void* stream);
int hcipy_fft_plan_destroy(hcipy_fft_plan plan);
/* Bytes of workspace (the column-blocked intermediate of the row pass) for `batch` fields in
 * either direction; 0 in 1-D. */
int hcipy_fft_workspace_size(hcipy_fft_plan plan, int64_t batch, size_t* bytes);
/* direction 0: forward, alpha * sum_n f[n] exp(-i x_n . u_k) on the output grid; 1: backward,
 * alpha * sum_k F[k] exp(+i u_k . x_n) on the input grid.  The reference's weights are the
 * caller's alpha.  `batch` fields of the plan's dtype, contiguous; in and out must not alias. */
int hcipy_fft_execute(hcipy_fft_plan plan, int direction, const void* in, void* out, int64_t batch, double alpha_re,
                      double alpha_im, void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HCIPY_AMD_H_ */
