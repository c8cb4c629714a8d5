/*
 * ocx - MI355X (gfx950) tracer advection of pyOCEAN: WENO5-Z/JS faces, Zalesak FCT, SSP-RK3, on the masked
 * C-grid.  C ABI of libocx.so.
 *
 * Conventions
 *   - All array pointers are DEVICE pointers (HBM), C-contiguous, FP64 unless stated, index order
 *     (tracer t, layer k, row j, column i), k = 0 at the bed.
 *     stacks (nt, nz, ny, nx); cells V0, V1, ok (nz, ny, nx), ok one byte 0/1; faces Qx (nz, ny, nx+1),
 *     Qy (nz, ny+1, nx), Fz (nz+1, ny, nx); face stacks fx (nt, nz, ny, nx+1), fy (nt, nz, ny+1, nx).
 *   - stream: a hipStream_t (NULL = the default stream).  Calls are asynchronous on it.
 *   - Every call returns 0 on success or a negative OCX_E_* code; ocx_last_error() then describes it
 *     (thread-local).  Shapes and pointers are checked on the host before any launch; kernels never read or write
 *     outside the extents implied by the shapes.
 *   - Results are bit-identical to the reference's FP64 arithmetic (no contraction, IEEE divisions) and do not
 *     depend on launch configuration.
 */
#ifndef OCX_H
#define OCX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* flags */
#define OCX_OPEN_W 1u      /* west edge open (upstream values on its faces) */
#define OCX_OPEN_E 2u
#define OCX_OPEN_S 4u
#define OCX_OPEN_N 8u
#define OCX_WENO_JS 16u    /* WENO5-JS instead of WENO5-Z (tracer_scheme == "weno5js") */
#define OCX_UNLIMITED 32u  /* FCT coefficients = 1: the unlimited RK3 result (advect_fct(unlimited=True)) */
#define OCX_UP3 64u        /* UP3 faces (tracer_scheme == "upwind3"): RK3 step, RHS and faces_h only */
#define OCX_UP1 128u       /* first-order upwind faces (tracer_scheme == "upwind1"): as OCX_UP3 */
#define OCX_CENTRED 256u   /* centred 6th/4th/2nd-order faces (rhs/flux_x with centred=True): RHS and faces_h only */

/* error codes */
#define OCX_OK 0
#define OCX_E_ARG (-1)        /* invalid shape, pointer or flag */
#define OCX_E_HIP (-2)        /* a HIP runtime call failed */
#define OCX_E_WORKSPACE (-3)  /* workspace missing or too small */
#define OCX_E_RANGE (-4)      /* the inputs violate a bound the reference raises on (an emptied layer, CFL) */

const char* ocx_last_error(void);
int ocx_version(void);

/* Device context (the survey's proposed ocx_create / ocx_alloc / ocx_h2d boundary; no counterpart in the reference,
 * whose arrays live in host NumPy): a device and a stream on it (default priority), for a host that does not link
 * HIP itself.  Pass ocx_ctx_stream(ctx) as the `stream` of the compute calls and pointers from ocx_alloc as their
 * arrays.  ocx_h2d / ocx_d2h copy on the context's stream and return when the copy is complete; ocx_free waits
 * for the stream first.  A context is used by one host thread at a time. */
typedef struct ocx_ctx ocx_ctx;
int ocx_create(int device, ocx_ctx** out);
int ocx_destroy(ocx_ctx* ctx);
void* ocx_ctx_stream(ocx_ctx* ctx);
int ocx_alloc(ocx_ctx* ctx, size_t bytes, void** dptr);
int ocx_free(ocx_ctx* ctx, void* dptr);
int ocx_h2d(ocx_ctx* ctx, void* dst, const void* src, size_t bytes);
int ocx_d2h(ocx_ctx* ctx, void* dst, const void* src, size_t bytes);
int ocx_sync(ocx_ctx* ctx);

/* Workspace (device bytes) of ocx_advect_fct_f64 / ocx_advect_fct_split / ocx_advect_rk3_f64 for one call of that
 * shape. */
size_t ocx_fct_workspace_bytes(int nt, int nz, int ny, int nx);
size_t ocx_fct_split_workspace_bytes(int nt, int nz, int ny, int nx);
size_t ocx_rk3_workspace_bytes(int nt, int nz, int ny, int nx);

/* One SSP-RK3 + Zalesak FCT step of the stack C.
 * Replaces TracerAdvection.advect_fct (core/components/tracers.py:208-254) and its kernels _advect_faces_h,
 * _fct_stage, _fct_fluxes, _fct_coeff_kernel, _fct_final (core/dynamics/sigma_baroclinic.py:265-321, 407-530).
 * inflow: exterior values (nt, nz, ny, nx) or NULL (stage 1 and the donor flux then use C, stages 2 and 3 their
 * own values, tracers.py:231).  flux_face_i: -1, or a u-face column 0..nx whose effective flux F_L + C A is
 * written to fx_i (nt, nz, ny).  open_flux: NULL or (nt) net open-edge inflow rate per tracer (tracers.py:65-78).
 * C_out must not alias C. */
int ocx_advect_fct_f64(const double* C, const double* V0, const double* V1, const double* Qx, const double* Qy,
                       const uint8_t* ok, const double* inflow, int nt, int nz, int ny, int nx, double dt,
                       unsigned flags, int flux_face_i, double* C_out, double* fx_i, double* open_flux,
                       void* workspace, size_t workspace_bytes, void* stream);

/* The same step in split precision (no counterpart in the reference; the precision form of the same algorithm):
 * the WENO reconstruction (smoothness indicators, weights, face values and fluxes) in FP32, the stage values c1, c2
 * and the face fluxes F0, F0 + F1, A stored in FP32; the h*T accumulation, the low-order solution, the Zalesak
 * ratios, the limited update and content sums in FP64.  Conservation and the FCT bounds hold as in FP64; values
 * differ from the FP64 step by the FP32 reconstruction error.  Same arguments as ocx_advect_fct_f64 except that
 * OCX_UNLIMITED is rejected. */
int ocx_advect_fct_split(const double* C, const double* V0, const double* V1, const double* Qx, const double* Qy,
                         const uint8_t* ok, const double* inflow, int nt, int nz, int ny, int nx, double dt,
                         unsigned flags, int flux_face_i, double* C_out, double* fx_i, double* open_flux,
                         void* workspace, size_t workspace_bytes, void* stream);

/* The FP64 step of ocx_advect_fct_f64 for callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc); no counterpart in the reference.  Same arguments, meaning, results and
 * errors as ocx_advect_fct_f64, without the workspace: the step's workspace is a device array inside the library of
 * OCX_NOALLOC_SCRATCH_BYTES (64 MiB, held on each GPU the library runs on until the process ends), so a shape whose
 * ocx_fct_workspace_bytes exceeds it is rejected with OCX_E_WORKSPACE (e.g. 2 x 20 x 128 x 128 needs 32 MB and
 * fits).  The call allocates nothing (no hipMalloc, hipFree or allocating library routine).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other host threads, other streams not ordered after this
 * one) race and must be serialised by the caller. */
#define OCX_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int ocx_advect_fct_f64_noalloc(const double* C, const double* V0, const double* V1, const double* Qx,
                               const double* Qy, const uint8_t* ok, const double* inflow, int nt, int nz, int ny,
                               int nx, double dt, unsigned flags, int flux_face_i, double* C_out, double* fx_i,
                               double* open_flux, void* stream);

/* One SSP-RK3 step without limiter, x, y and (if vert) z sweeps.  Replaces TracerAdvection.advect
 * (core/components/tracers.py:342-375) with _advect_div and _rk_stage (sigma_baroclinic.py:324-364, 856-875).
 * Fz (nz+1, ny, nx) may be NULL when vert == 0; vert as the reference's any(Fz != 0).  fx_i as above
 * ((F(c0) + F(c1))/6 + 2 F(c2)/3 of column flux_face_i).  flags may carry OCX_UP3 or OCX_UP1 (not with
 * OCX_WENO_JS): the faces of rhs_numpy (tracers.py:139-151, _up3_faces sigma_baroclinic.py:164-187) instead of
 * WENO, whose z faces are always formed, so Fz is then required and vert ignored. */
int ocx_advect_rk3_f64(const double* C, const double* V0, const double* V1, const double* Qx, const double* Qy,
                       const double* Fz, const uint8_t* ok, const double* inflow, int nt, int nz, int ny, int nx,
                       double dt, unsigned flags, int vert, int flux_face_i, double* C_out, double* fx_i,
                       void* workspace, size_t workspace_bytes, void* stream);

/* The 'pc' step of TracerAdvection.advect with tracer_time == "pc" (advect_numpy, tracers.py:303-330): a half step
 * with centred faces on x, y and z, ch = (V0 C + dt/2 r_c(C)) / (V0 + dV/2), then the full step from C with the
 * faces of the flags (WENO5-Z, OCX_WENO_JS, OCX_UP3 or OCX_UP1) of ch, out = (V0 C + dt r(ch)) / V1; both
 * divisions fall back to C where the volume is not > 0.  Fz is required (the predictor always forms the z faces);
 * vert as in ocx_advect_rk3_f64 for the WENO corrector.  fx_i: the corrector's x-face column flux_face_i, F(ch).
 * Arguments, workspace (ocx_rk3_workspace_bytes) and errors as ocx_advect_rk3_f64. */
int ocx_advect_pc_f64(const double* C, const double* V0, const double* V1, const double* Qx, const double* Qy,
                      const double* Fz, const uint8_t* ok, const double* inflow, int nt, int nz, int ny, int nx,
                      double dt, unsigned flags, int vert, int flux_face_i, double* C_out, double* fx_i,
                      void* workspace, size_t workspace_bytes, void* stream);

/* WENO horizontal face fluxes (fx, fy) of the stack c.  Replaces _advect_faces_h (sigma_baroclinic.py:265-321)
 * as called by TracerAdvection.faces_h (tracers.py:166-176).  ext: exterior values or NULL (= c).  With OCX_UP3 /
 * OCX_UP1 / OCX_CENTRED: the face fluxes that rhs_numpy differences (flux_x for those schemes or centred=True,
 * tracers.py:194-206; _centred_faces sigma_baroclinic.py:632-652). */
int ocx_faces_h_f64(const double* c, const double* Qx, const double* Qy, const uint8_t* ok, const double* ext,
                    int nt, int nz, int ny, int nx, unsigned flags, double* fx, double* fy, void* stream);

/* Face VALUES (not fluxes) along one axis of any array, as TracerAdvection.faces (core/components/tracers.py:110-115)
 * returns them: _weno5_faces (WENO5-Z, or WENO5-JS with OCX_WENO_JS; sigma_baroclinic.py:564-578 with _weno5_rows
 * :236-262), _up3_faces (OCX_UP3 / OCX_UP1; :164-187) or _centred_faces (OCX_CENTRED, flux unused and may be NULL;
 * :632-652).  The arrays are viewed as c, ok (outer, n, inner) and flux, out (outer, nf, inner), the face axis in
 * the middle (any axis of a C-contiguous array: outer = the product of the dimensions before it, inner = of those
 * after it).  Face f lies between cells f-1 and f; faces 1 .. n-1 take the scheme's value, upwind-ordered by the
 * sign of flux (flux >= 0: the low side; NaN: the high side), cells outside 0 .. n-1 invalid; all other faces 0. */
int ocx_faces_f64(const double* c, const double* flux, const uint8_t* ok, long long outer, int n, int nf,
                  long long inner, unsigned flags, double* out, void* stream);

/* Open-edge face values of one horizontal axis, in place: face (nt, nz, ny, nx+1) with Q (nz, ny, nx+1) for axis 0
 * (west and east edges), face (nt, nz, ny+1, nx) with Q (nz, ny+1, nx) for axis 1 (south and north).  On each
 * edge in open_mask (OCX_OPEN_*) the face takes ext (nt, nz, ny, nx; NULL = c) of its boundary cell where the flux
 * enters (Q > 0 west / south, Q < 0 east / north), else the boundary cell of c.  Replaces
 * TracerAdvection.open_faces (core/components/tracers.py:117-137). */
int ocx_open_faces_f64(double* face, const double* c, const double* Q, const double* ext, int nt, int nz, int ny,
                       int nx, int axis, unsigned open_mask, void* stream);

/* Donor-cell face fluxes.  Replaces TracerAdvection.upwind_faces_h (tracers.py:178-193). */
int ocx_upwind_faces_h_f64(const double* c, const double* Qx, const double* Qy, const double* ext, int nt, int nz,
                           int ny, int nx, unsigned flags, double* fx, double* fy, void* stream);

/* Advection RHS -div(Q c) with x, y and (vert) z WENO faces.  Replaces _advect_div (sigma_baroclinic.py:324-364)
 * as called by TracerAdvection.rhs (tracers.py:153-164); with OCX_UP3 / OCX_UP1 replaces rhs_numpy
 * (tracers.py:139-151): x, y and z UP faces, Fz required; OCX_CENTRED the same with centred faces
 * (rhs(centred=True)). */
int ocx_rhs_f64(const double* c, const double* Qx, const double* Qy, const double* Fz, const uint8_t* ok,
                const double* ext, int nt, int nz, int ny, int nx, unsigned flags, int vert, double* out,
                void* stream);

/* Zalesak coefficients (Cx, Cy) of antidiffusive fluxes (Ax, Ay).  Replaces _fct_coeff_kernel
 * (sigma_baroclinic.py:407-449).  workspace: 2 * nt*nz*ny*nx doubles. */
int ocx_fct_coefficients_f64(const double* c0, const double* cL, const double* V1, const double* Ax,
                             const double* Ay, const uint8_t* ok, double dt, int nt, int nz, int ny, int nx,
                             double* Cx, double* Cy, void* workspace, size_t workspace_bytes, void* stream);

/* Periodic pad/crop (Periodic.pad / crop, core/model/periodic.py:71-95) of `lead` slabs of a (sy, sx) array.
 * kind: 0 cells, 1 u-faces (x staggered), 2 v-faces (y staggered).  pad_x/pad_y: halo on that axis (0 = axis not
 * periodic).  elem: 8 (FP64) or 1 (bytes).  crop != 0: the inverse (dst gets the interior of src). */
int ocx_periodic_pad(const void* src, void* dst, size_t lead, int sy, int sx, int kind, int pad_y, int pad_x,
                     int elem, int crop, void* stream);

/* Explicit Laplacian diffusion along the layers of an ALE step (_diffuse_ale_kernel,
 * core/dynamics/sigma_baroclinic.py:533-561, as TracerDiffusion.diffuse_h calls it, core/components/tracers.py:420-446),
 * same per-element arithmetic, for a stack of nt tracers: T, Tw (nt, nz, ny, nx) the field and the field diffused
 * (the anomaly or T itself); h (nz, ny, nx) layer thickness already zeroed on inactive cells; act (ny, nx);
 * cx = L_u/dx_u (ny, nx+1), cy = L_v/dy_v (ny+1, nx); okx (ny, nx+1), oky (ny+1, nx) face masks (bytes);
 * V (nz, ny, nx) volumes.  out (nt, nz, ny, nx) must not alias T or Tw. */
int ocx_diffuse_h_f64(const double* T, const double* Tw, const double* h, const uint8_t* act, const double* cx,
                      const double* cy, const uint8_t* okx, const uint8_t* oky, const double* V, int nt, int nz, int ny,
                      int nx, double kh, double dt, double* out, void* stream);

/* Implicit vertical diffusion of nt stacked fields (vertical_implicit / _thomas_columns,
 * core/dynamics/sigma_wd.py:57-99), same arithmetic: u, out (nt, n, columns) (out may be u); dz (n, columns);
 * nu (n+1, columns) interface diffusivities; r_bed NULL, (columns) on layer 0 (r_layers = 1) or (n, columns)
 * (r_layers = n); tau_top NULL or (columns) surface flux into the top layer.  workspace:
 * ocx_vertical_implicit_workspace_bytes(n, columns) bytes (the forward-sweep coefficients, and for n <= 64 each
 * level's lower coefficient and pivot, shared by the fields). */
size_t ocx_vertical_implicit_workspace_bytes(int n, long long columns);
int ocx_vertical_implicit_f64(const double* u, const double* dz, const double* nu, const double* r_bed, int r_layers,
                              const double* tau_top, int nt, int n, long long columns, double dt, double* out,
                              void* workspace, size_t workspace_bytes, void* stream);

/* Numerical mixing of the advection, out = V1 (<c^2> - c^2) / dt per cell (accumulate != 0: out += it), with c and
 * c^2 (nz, ny, nx) two members of the advected stack (T and T^2).  Replaces mix_num (core/model/driver.py:905) and
 * VarianceBudget.advective (core/components/variance_budget.py:60-62), same operation order. */
int ocx_mix_num_f64(const double* c, const double* c2, const double* V1, int nz, int ny, int nx, double dt,
                    int accumulate, double* out, void* stream);

/* Per-tracer content sum_k,j,i V c (FP64, compensated, in a fixed order that does not depend on the launch):
 * out (nt).  The global content sums of the reference's conservation checks (tracers.py:359-375 callers,
 * test_sigma_ale.py:1342-1361).  workspace: ocx_content_workspace_bytes(nt) bytes of device memory. */
size_t ocx_content_workspace_bytes(int nt);
int ocx_content_f64(const double* c, const double* V, int nt, int nz, int ny, int nx, double* out, void* workspace,
                    size_t workspace_bytes, void* stream);

/* Self-test of the correctly rounded fast divisions the FP64 kernels use in place of the IEEE division (no
 * counterpart in the reference, whose Numba kernels divide with IEEE semantics, sigma_baroclinic.py:212-231): for n
 * device operand pairs (x[e], y[e]), out (6, n) receives div_const(x, 6), div_const(x, 3) (the divisions by the
 * WENO and F_H constants, defined for finite x), div_nr(x, y) (reciprocal, two Newton steps and one residual correction; defined for
 * y > 0 with both operands inside 2^-500 .. 2^500 or x = +-0), qdiv(x, y) (div_nr inside that range, IEEE outside),
 * the device's IEEE quotient x / y, and qdiv_r(x, recip_of(y)) (qdiv with the reciprocal of y refined ahead of x).
 * Each of the fast forms equals the IEEE quotient bit for bit in its stated range (qdiv and qdiv_r everywhere). */
int ocx_selftest_div_f64(const double* x, const double* y, long long n, double* out, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Batched form (no counterpart in the reference, which advances one domain per call): many independent domains
 * (ensemble members or regional domains of uneven nt, nz, ny, nx, each FP64 or split precision) advanced by one
 * FCT step with one persistent launch per stage (four launches in all), optionally replayed as a captured hipGraph.
 * Each domain's result is the one ocx_advect_fct_f64 / ocx_advect_fct_split gives alone (same kernels, same
 * per-domain arithmetic).  The state ping-pongs: parity 0 reads C and writes C_out, parity 1 the reverse.
 * No inflow-less restriction; no fx_i / open_flux side outputs. */
typedef struct ocx_domain {
    double* C;                    /* (nt, nz, ny, nx), state buffer A */
    double* C_out;                /* (nt, nz, ny, nx), state buffer B */
    const double* V0;
    const double* V1;
    const double* Qx;
    const double* Qy;
    const double* inflow;         /* NULL: no exterior values */
    const uint8_t* ok;
    int nt, nz, ny, nx;
    double dt;
    unsigned flags;               /* OCX_OPEN_*, OCX_WENO_JS */
    int precision;                /* 0: FP64 (bit-identical to the reference), 1: split */
    void* workspace;              /* ocx_fct_workspace_bytes / ocx_fct_split_workspace_bytes of the domain */
    size_t workspace_bytes;
} ocx_domain;

typedef struct ocx_batch ocx_batch;

/* rows_per_chunk: rows one wavefront marches (0: 64); waves_per_group: wavefronts per workgroup, the tracers of one
 * layer (0: 2, at most 4); groups_per_cu: resident workgroups per compute unit of the persistent grid (0: 4).  Copies
 * the domain table to the device (synchronous). */
int ocx_batch_create(const ocx_domain* domains, int ndomains, int rows_per_chunk, int waves_per_group,
                     int groups_per_cu, ocx_batch** out);
int ocx_batch_step(ocx_batch* batch, int parity, void* stream);
/* Captures the step of this parity into a hipGraph on its first call (stream must not be the NULL stream), then
 * launches the graph. */
int ocx_batch_graph_step(ocx_batch* batch, int parity, void* stream);
int ocx_batch_destroy(ocx_batch* batch);

/* Launch shape of the single-domain entry points (for the autotuner; no counterpart in the reference):
 * rows_per_chunk and waves_per_group as for ocx_batch_create; 0 restores the default.  Process-wide. */
int ocx_set_tuning(int rows_per_chunk, int waves_per_group);

/* Per-kernel timing (no counterpart in the reference; for benchmarks): while enabled, every kernel launch of this
 * library is bracketed by a hipEvent pair on its stream.  ocx_profile_collect waits for the recorded events,
 * returns the number of distinct kernels and fills up to max_kernels entries (static name strings, total
 * milliseconds, launch counts), then forgets the records. */
int ocx_profile_enable(int on);
int ocx_profile_collect(int max_kernels, const char** names, double* ms_total, long long* launches);

/* ---------------------------------------------------------------------------------------- ALE layer transport
 * The producer of the advection step's inputs (Qx, Qy, V0, V1, ok) in the ALE vertical coordinate:
 * ALELayerTransport.__call__ (core/components/layer_transport.py:111-156) without open boundaries and freshwater.
 * Face arrays of one direction in their natural layout: axis 0 the u-faces (n, ny, nx+1) with cells along x,
 * axis 1 the v-faces (n, ny+1, nx) with cells along y (the reference calls the same functions on swapped axes;
 * its (n, nx, ny+1) arrays are the transposes of these).  Cells (n, ny, nx); 2-D cell fields (ny, nx); active one
 * byte 0/1.  q_pos is the transport thickness for flow towards + (from the low-index cell, the reference's q_e /
 * q_n), q_neg for flow towards - (q_w / q_s). */

/* u^θ and Q_bt of one face direction, columns = the face count of one layer.
 * Replaces transport_velocity / _transport_velocity_kernel (core/vertical/ale.py:789-829).  theta_bc: pass theta
 * for the reference's theta_bc=None. */
int ocx_transport_velocity_f64(const double* u_new, const double* u_old, const uint8_t* last_wet, const double* w,
                               const double* L, const double* D, int n, long long columns, double theta,
                               double theta_bc, double* ut, double* Q_bt, void* stream);

/* Transport thickness at the faces of one axis; clip: cut at the shallower bed (H the rest depth, NULL without
 * clip).  Replaces face_transport_thickness / _ftt_kernel (core/vertical/ale.py:168-230). */
int ocx_face_transport_thickness_f64(const double* h, const double* H, const uint8_t* active, int n, int ny, int nx,
                                     int axis, int clip, double* q_pos, double* q_neg, void* stream);

/* PPM edge thicknesses along one axis (e_lo: west / south edge, e_hi: east / north), cells (n, ny, nx).
 * Replaces ppm_edges_x / _ppm_edges_kernel (core/vertical/ale.py:440-476, 578-623). */
int ocx_ppm_edges_f64(const double* h, const uint8_t* active, int n, int ny, int nx, int axis, double h_min,
                      double* e_lo, double* e_hi, void* stream);

/* Layer volume fluxes with PPM thickness and the Newton correction per face column (all columns iterate until all
 * have converged, at most max_iter passes), then the residual share.  d_t: the cell width along the axis (ny, nx);
 * L, Q_bt: face arrays of one layer.  Replaces ppm_transport_x / _ppm_transport_kernel (core/vertical/ale.py:626-764).
 * ocx_ppm_transport_iterations reads back (synchronously) how many Newton passes the last call on this workspace
 * took. */
size_t ocx_ppm_transport_workspace_bytes(int n, int ny, int nx, int axis, int max_iter);
int ocx_ppm_transport_f64(const double* h, const double* ut, const double* L, const double* Q_bt, const double* q_pos,
                          const double* q_neg, const uint8_t* active, const double* d_t, int n, int ny, int nx,
                          int axis, double dt, int max_iter, double* out, void* workspace, size_t workspace_bytes,
                          void* stream);
int ocx_ppm_transport_iterations(const void* workspace, int n, int ny, int nx, int axis, int max_iter, void* stream);

/* Layer volume fluxes with the donor-cell (upwind) thickness, corrected to the free-surface flux Q_bt: the
 * thickness_flux "upwind" option of the layer transport.  ut, q_pos, q_neg, out: face stacks of the axis in their
 * natural layout; L, Q_bt: one layer of faces.  Replaces upwind_transport_x (core/vertical/ale.py:233-251). */
int ocx_upwind_transport_f64(const double* ut, const double* L, const double* Q_bt, const double* q_pos,
                             const double* q_neg, int n, int ny, int nx, int axis, double* out, void* stream);

/* Layer fluxes through the open edges' faces, in place on Qx (n, ny, nx+1) and Qy (n, ny+1, nx): on every edge of
 * open_mask (1 west, 2 east, 4 south, 8 north) and face with free-surface face depth open_D > 0, Q_k = (h_k / sum h)
 * Q_bt + L h_k (theta_b u'_new + (1 - theta_b) u'_old) of the interior cell's thicknesses h (n, ny, nx), with
 * Q_bt = open_Q (the Flather flux); theta_b the baroclinic time weight (theta when the model gives none).
 * Replaces OpenBoundary.open_face_transports (core/components/open_boundary.py:334-351) and open_face_transport
 * (core/vertical/ale.py:841-861). */
int ocx_open_face_transports_f64(double* Qx, double* Qy, const double* h, const double* u_old, const double* v_old,
                                 const double* u_new, const double* v_new, const double* open_Du,
                                 const double* open_Qu, const double* open_Dv, const double* open_Qv, const double* Lu,
                                 const double* Lv, int n, int ny, int nx, unsigned open_mask, double theta_b,
                                 void* stream);

/* Freshwater volumes: dv_scale dV (ny, nx) joins the top layer (k = n-1) of V (n, ny, nx) into out (dv_scale = dt
 * with dV = fw gives the layer transport's dt fw); a removal that would take the top layer below the floor
 * (floor_cols (ny, nx), or NULL for the scalar floor) is taken from the layers below, top-down, each first kept at
 * the floor, then down to zero.  With area (ny, nx), V_out = out A as well.  n_limited (host, optional: a host
 * synchronisation) counts those columns.  Replaces freshwater_volumes (core/components/freshwater.py:31-58) and
 * the layer transport's h_fw, V_fw (core/components/layer_transport.py:148-155).
 * workspace: ocx_freshwater_workspace_bytes(). */
size_t ocx_freshwater_workspace_bytes(void);
int ocx_freshwater_volumes_f64(const double* V, const double* dV, double dv_scale, const double* floor_cols,
                               double floor, const double* area, int n, int ny, int nx, double* out, double* V_out,
                               void* workspace, size_t workspace_bytes, long long* n_limited, void* stream);

/* Freshwater substep on the tracer stack (nt, n, ny, nx), in place: every member but T (t_index) and T^2 (t2_index,
 * -1 if absent) keeps its content, c V_adv / V_fw, where fw_area_dt (ny, nx) is nonzero; columns whose volume
 * changed below the top layer share one concentration over the touched layers.  var_surf (n, ny, nx, optional)
 * the salinity-variance source of member s_index.  n_limited (host, optional) counts the columns with a change
 * below the top layer.  Replaces apply_freshwater (core/components/freshwater.py:61-122). */
int ocx_apply_freshwater_f64(double* stack, int nt, int n, int ny, int nx, int t_index, int t2_index, int s_index,
                             const double* V_adv, const double* V_fw, const double* fw_area_dt, double dt,
                             double* var_surf, void* workspace, size_t workspace_bytes, long long* n_limited,
                             void* stream);

/* Outflow limiter, in place on Qx and Qy; host-synchronous (the number of passes depends on the data).
 * Replaces limit_outflow / _give_back / _outflow_exceeds (core/vertical/ale.py:299-437).  check_only: only test
 * whether any active cell-layer sends out more than it holds.  status: OCX_OUTFLOW_*; n_limited: cell-layers scaled
 * over all passes.  On OCX_OUTFLOW_EMPTY_DONOR where = (axis, a, f) of the face; on OCX_OUTFLOW_ACCEPTED /
 * NEGATIVE where = (k, j, i) of the lowest h and value = that h. */
#define OCX_OUTFLOW_NOOP 0          /* nothing exceeded: Qx, Qy untouched */
#define OCX_OUTFLOW_OK 1            /* limited; no cell-layer exceeds any more */
#define OCX_OUTFLOW_ACCEPTED 2      /* still exceeding after max_pass but h >= -h_floor (the reference warns) */
#define OCX_OUTFLOW_EXCEEDS 3       /* check_only: some cell-layer exceeds */
#define OCX_OUTFLOW_EMPTY_DONOR 4   /* returned with OCX_E_RANGE: a cut has nowhere to go */
#define OCX_OUTFLOW_NEGATIVE 5      /* returned with OCX_E_RANGE: h would fall below -h_floor */
size_t ocx_limit_outflow_workspace_bytes(int n, int ny, int nx);
int ocx_limit_outflow_f64(const double* h, double* Qx, double* Qy, const double* area, const uint8_t* active, int n,
                          int ny, int nx, double dt, double h_floor, int max_pass, int periodic_x, int periodic_y,
                          int check_only, void* workspace, size_t workspace_bytes, long long* n_limited, int* status,
                          long long* where, double* value, void* stream);

/* Lagrangian thickness h† = max(h - dt div / A, 0) with div from divq (n, ny, nx) or from Qx, Qy (then divq NULL);
 * optionally V0 = h A, V1 = h† A, ok = active & V1 > 0.  With bad_index non-NULL the call is host-synchronous and
 * returns OCX_E_RANGE with the first cell-layer below -tol (linear index, else -1).
 * Replaces lagrangian_thickness (core/vertical/ale.py:278-295) and the volume lines of
 * ALELayerTransport.__call__ (core/components/layer_transport.py:142-144). */
size_t ocx_lagrangian_workspace_bytes(void);
int ocx_lagrangian_thickness_f64(const double* h, const double* divq, const double* Qx, const double* Qy,
                                 const double* area, const uint8_t* active, int n, int ny, int nx, double dt,
                                 double tol, double* out, double* V0, double* V1, uint8_t* ok, void* workspace,
                                 size_t workspace_bytes, long long* bad_index, void* stream);

/* ---------------------------------------------------------------------------------------- vertical regrid + remap
 * ALECoordinate.regrid_remap (core/components/vertical_coordinate.py:69-79) for the z* and sigma rules.  Columns
 * are the (ny, nx) cells flattened: level arrays (n, columns), C-contiguous. */

/* Conservative remap of every wet column of the fields f (nf, n0, columns) from thicknesses h_src (n0, columns) to
 * h_tgt (n1, columns); scheme 0 PCM, 1 PLM, 2 PPM_H4 (lowered to PLM / PCM on columns with <= 3 / 1 active cells).
 * Dry columns (wet 0): f copied when n0 == n1, else NaN.  out (nf, n1, columns).
 * Replaces remap_columns / _remap_columns / remap_column (core/vertical/remap.py:259-427). */
int ocx_remap_columns_f64(const double* h_src, const double* f, const double* h_tgt, const uint8_t* wet, int n0,
                          int n1, long long columns, int nf, int scheme, double* out, void* stream);

/* New layer thicknesses of the z* (rule 0, dz the nominal thicknesses surface first) or sigma (rule 1, dz the
 * fractions bed first) rule: wet = active & H + eta > 0 & sum h > 0 (H with NaN read as 0, written to wet if
 * non-NULL); h_new = the rule's layers where wet, h elsewhere.  With bad_column non-NULL the call is
 * host-synchronous and returns OCX_E_RANGE naming the first wet column the rule cannot build (the reference raises
 * ValueError there).  Replaces ALECoordinate.regrid (vertical_coordinate.py:94-109) with regrid_columns and
 * zstar_column / sigma_column / enforce_min_thickness (core/vertical/regrid.py:41-95, 325-405). */
size_t ocx_regrid_workspace_bytes(void);
int ocx_regrid_f64(const double* H, const double* eta, const uint8_t* active, const double* h, int n,
                   long long columns, int rule, const double* dz, double h_min, double* h_new, uint8_t* wet,
                   void* workspace, size_t workspace_bytes, long long* bad_column, void* stream);

/* The hybrid (HYCOM1) rule: interfaces on the target densities rho_target (n+1, surface first) of the layer
 * densities rho (n, columns) by P1M_H2 interpolation, isopycnal layers kept in place (keep), the z* floor dz_top
 * (surface first), the maximum layer thicknesses max_h_top, the shallow sigma window between d_sigma and d_hybrid
 * (d_hybrid 0: none; dsigma bed first) and the minimum thickness h_min.  wet, h_new and the bad-column report as
 * ocx_regrid_f64; the workspace holds the columns' working arrays ((6 n + 3) doubles per column).
 * Replaces regrid_columns('hybrid') / hycom1_column / _p1m_h2_positions / _keep_consistent / window_shallow_sigma
 * (core/vertical/regrid.py:98-357) under ALECoordinate.regrid (vertical_coordinate.py:94-109). */
size_t ocx_regrid_hybrid_workspace_bytes(int n, long long columns);
int ocx_regrid_hybrid_f64(const double* H, const double* eta, const uint8_t* active, const double* h, const double* rho,
                          int n, long long columns, const double* dz_top, const double* dsigma,
                          const double* rho_target, const double* max_h_top, double d_sigma, double d_hybrid,
                          double h_min, int keep, double* h_new, uint8_t* wet, void* workspace,
                          size_t workspace_bytes, long long* bad_column, void* stream);

/* Layer densities of the hybrid rule, ALECoordinate.regrid_density (core/components/vertical_coordinate.py:80-100),
 * elementwise over n cells: eos 0 = the linear EOS, rho0 - t_coef (T - t_ref) [+ s_coef (S - s_ref)]
 * (LinearEOS.density, core/components/eos.py:62-69; linear = host array {rho0, t_coef, t_ref, s_coef, s_ref},
 * S may be NULL); eos 1 = Wright's full-range density at sea pressure p [Pa] (WrightFullEOS.density,
 * physics/eos/wright_full.py:80-89; S required; p a device array of n values or NULL for p_scalar at every cell,
 * e.g. the reference pressure 2e7 Pa of the potential density).  T, S, p, rho are device arrays. */
int ocx_regrid_density_f64(int eos, const double* T, const double* S, const double* p, double p_scalar,
                           const double* linear, long long n, double* rho, void* stream);

#ifdef __cplusplus
}
#endif
#endif
