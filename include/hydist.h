/*
 * hydist.h -- C ABI of libhydist.so, the MI355X (gfx950) FP64 ADI time step of
 * the HYDIST 2-D shallow-water and suspended-sediment model.
 *
 * The reference drives its CUDA kernels from Python through PyCUDA:
 * SourceModule(...).get_function(name) and one launch per kernel
 * (Engine.py:73-128, Engine.py:144-366, rep.py:115-139), with arrays owned by
 * Pointers (Pointers.py:15-166) and packed into the device structs
 * Argument_Pointers / Array_Pointers (constant.cuh:6-29, Pointers.py:169-191).
 * This library replaces that boundary: the Python host (hydist_amd package)
 * binds these functions with ctypes.  No torch or HIP types cross it: plain
 * pointers, sizes and the two POD structs below.
 *
 * Conventions (identical to the reference's, Global_Variables.py:68-127):
 *  - N rows (index i, x, u, UZ lines run along i), M columns (index j, y, v,
 *    VZ lines run along j); every 2-D field is (N+3) x (M+3) row-major,
 *    flat index i*(M+3)+j, FP64 except the int32 mask khouot.
 *  - Segment tables daui/cuoii [(N+2) x SL], dauj/cuoij [(M+2) x SL],
 *    moci [N+2], mocj [M+2] (int32), SL = segment_limit.
 *  - Boundary series bc_* / CC_* are laid out seg*total_time + hour.
 *
 * Ownership: the library owns all device memory; the caller owns host buffers,
 * which are copied synchronously by hd_upload / hd_download.  Errors: every
 * function returns HD_OK (0) or a negative HD_E* code and never aborts;
 * hd_last_error() gives the message.  Threading: one hd_ctx per host thread;
 * all work is queued on the context's own HIP stream.
 */
#ifndef HYDIST_H
#define HYDIST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HD_OK 0
#define HD_EINVAL (-1)        /* bad argument */
#define HD_ESHAPE (-2)        /* byte count does not match the field layout */
#define HD_ENAME (-3)         /* unknown field name */
#define HD_EHIP (-4)          /* HIP runtime error */
#define HD_ESEGOVERFLOW (-5)  /* a line has more wet runs than segment_limit */
#define HD_ESTATE (-6)        /* call out of order (e.g. step before init) */
#define HD_ERANGE (-7)        /* boundary series too short for the time t */
#define HD_ENOMEM (-8)

/* Case parameters: the __constant__ values of a constant.cuh case header
 * (constant.cuh:31-118, PI from constant.cuh:3), evaluated in IEEE double
 * exactly as written there.  The last five members are the header's constant
 * device functions Windx/Windy/Ufr/Dxr/wss (constant.cuh:126-133); hd_create
 * evaluates them from the other members and ignores the values passed in. */
typedef struct hd_params {
    int32_t segment_limit;
    int32_t pad0;
    double dX, dY, dT;
    double dXbp, dX2, dYbp, dY2;
    double dTchia2dX, dTchia2dY;
    double QuyDoiTime, QuyDoiPi, HaiChiadT;
    double kenhhepng, kenhhepd;
    double NANGDAY, H_TINH;
    double Wind, huonggio;
    double Zbandau, heso, mu_mn;
    double NDnen, NDbphai, NDbtrai, NDbtren, NDbduoi;
    double Tod, Toe, hstoe, ghtoe, Mbochat;
    double ro, ros, dm, d90, muy, Dorong, Sx;
    double sohn, soha, sohn1;
    double CORIOLIS_FORCE, g, Ks;
    double PI;
    double Windx, Windy, Ufr, Dxr, wss;
} hd_params;

/* Boundary handling, Engine.py:29-46: the reference forces the analytic tide
 * (boundary_up/down/left/right, support_funcs.cu:356-429) whenever a narrow
 * channel flag is set, and the hourly series (Update_Boundary_Value,
 * support_funcs.cu:622-668) otherwise. */
#define HD_BND_AUTO 0
#define HD_BND_SERIES 1
#define HD_BND_ANALYTIC 2

/* Bed-load flux rule: REFERENCE computes Qb only where VTH == 0, as
 * Sediment_Transport.cu:365 is written; FIXED computes it where VTH != 0. */
#define HD_QB_REFERENCE 0
#define HD_QB_FIXED 1

/* Bed-change trigger: PERIODIC fires when int(t) % bed_change_start == 0 at an
 * integral t (Engine.py:364); EVERY_STEP fires at every step with
 * t >= bed_change_start. */
#define HD_BED_PERIODIC 0
#define HD_BED_EVERY_STEP 1

/* Keyword options of hydraulic_Calculation (Engine.py:49-50) that change the
 * step, plus the explicit modes above. */
typedef struct hd_step_opts {
    int32_t kenhhepd;
    int32_t kenhhepng;
    int32_t boundary_mode;
    int32_t ketdinh;
    int32_t qb_mode;
    int32_t bed_mode;
    double sediment_start;
    double bed_change_start;
} hd_step_opts;

typedef struct hd_ctx hd_ctx;

/* Library / context */
const char* hd_version(void);
const char* hd_error_string(int code);
/* Create a context on HIP device `device` for an N x M case.  series_len is
 * the length (doubles) of each boundary series bc_* and CC_*; total_time is
 * the number of hourly samples per segment (Global_Variables.py:53). */
int hd_create(int device, const hd_params* p, int N, int M, int series_len,
              int total_time, hd_ctx** out);
int hd_destroy(hd_ctx* c);
const char* hd_last_error(const hd_ctx* c);
/* Evaluated case parameters (derived members filled in). */
int hd_get_params(const hd_ctx* c, hd_params* out);

/* Named fields: the Pointers keys of rep.py:62-70 (VISCOINDX is accepted as an
 * alias of VISCOIDX).  dtype: 0 = float64, 1 = int32. */
int hd_field_info(hd_ctx* c, const char* name, size_t* nbytes, int* dtype);
int hd_upload(hd_ctx* c, const char* name, const void* src, size_t nbytes);
int hd_download(hd_ctx* c, const char* name, void* dst, size_t nbytes);
/* hmax_u, hmax_d, hmax_l, hmax_r (Argument_Pointers, constant.cuh:8). */
int hd_get_hmax(hd_ctx* c, double out[4]);

/* One-time initialisation, rep.py:126-139: Onetime_init, the segment finders,
 * Htuongdoi and preprocess_data.  Call after uploading the input fields. */
int hd_init(hd_ctx* c);
/* rep.py:135-139 alone (segment finders, Htuongdoi, preprocess_data) on the
 * current state: after uploading a pick-up state (rep.py:43-55, 133-134). */
int hd_prepare(hd_ctx* c);

/* One full ADI step t -> t + dT (both half-steps, sediment and bed change as
 * the options select), Engine.py:144-366.  *t_out receives the new time. */
int hd_step(hd_ctx* c, double t, const hd_step_opts* o, double* t_out);
/* nsteps consecutive steps, replayed from a captured HIP graph when the
 * sequence is graph-capturable; identical results to nsteps hd_step calls. */
int hd_run(hd_ctx* c, double t, int nsteps, const hd_step_opts* o, double* t_out);
/* Stages of a step, in Engine.py order, for drivers that interleave host work.
 * which: 0 = UZ half-step (first half), 1 = VZ half-step (second half).
 * t is the time after the half-step's t += dT/2. */
int hd_halfstep(hd_ctx* c, int which, double t, const hd_step_opts* o);
int hd_bed_change(hd_ctx* c, double t, const hd_step_opts* o);

/* hd_run that also adds, after every step, the number of wet cells
 * (khouot == 0 over the padded grid) of the step's end state to a device
 * counter: the wet-cell-steps of the run's steps 2..n+1 start states. */
int hd_run_counted(hd_ctx* c, double t, int nsteps, const hd_step_opts* o, double* t_out);
int hd_reset_wet_counter(hd_ctx* c);
int hd_wet_cell_steps(hd_ctx* c, long long* out);

/* Number of wet cells (khouot == 0) over the padded grid. */
int hd_wet_count(hd_ctx* c, long long* out);
int hd_sync(hd_ctx* c);
/* Device addresses of the Argument_Pointers / Array_Pointers mirrors
 * (constant.cuh:6-29), for callers that hand the structs to their own code. */
int hd_struct_ptrs(hd_ctx* c, void** arg, void** arr);
/* Per-kernel timing with HIP events recorded on the context's stream around
 * every launch.  hd_profile(c, 1) clears the accumulators and turns it on
 * (hd_run then enqueues steps directly instead of replaying graphs);
 * hd_kernel_times fills up to n (name, total ms, launch count) entries in
 * first-launch order and returns how many it filled. */
int hd_profile(hd_ctx* c, int enable);
int hd_kernel_times(hd_ctx* c, int n, const char** names, double* total_ms, int* launches);

/* After a caller (or a surrogate) has written u, v, z: copy them into the
 * half-step buffers t_u, t_v, t_z, set H_moi = htaiz + z on wet cells and
 * recompute the face depths Htdu / Htdv (Htuongdoi, support_funcs.cu:338-354),
 * so that the sediment kernels see that flow.  The mask is unchanged. */
int hd_flow_refresh(hd_ctx* c);
/* The sediment half of nsteps steps (boundary values, VTH / hesoK, the FS
 * transport sweeps and FS update of each half-step, bed change when
 * triggered: Engine.py:213-239, 331-366) on the current, frozen flow. */
int hd_sediment_run(hd_ctx* c, double t, int nsteps, const hd_step_opts* o, double* t_out);

/* Bed-change extrapolation (morphological acceleration over a frozen flow):
 * hd_bed_snapshot keeps a copy of dH; hd_bed_extrapolate then sets
 * dH += factor * (dH - snapshot), i.e. repeats the bed change made since the
 * snapshot `factor` more times. */
int hd_bed_snapshot(hd_ctx* c);
int hd_bed_extrapolate(hd_ctx* c, double factor);

/* Water-level check of a step from a saved state: hd_z_snapshot keeps a copy
 * of z; hd_z_change returns, over the wet cells of [1,N]x[1,M], the largest
 * |z - z_saved| (a NaN counts as infinite) and the sums of htaiz + z now and
 * at the snapshot (the wet volume per unit cell area). */
int hd_z_snapshot(hd_ctx* c);
int hd_z_change(hd_ctx* c, double* max_abs_dz, double* wet_volume, double* wet_volume_snap);
/* hd_z_change with a bound: also the fraction of the wet cells whose
 * |z - z_saved| exceeds `bound` (a NaN counts as exceeding).  The largest
 * change sits at wetting and drying cells, where one plain step moves z by
 * up to about a centimetre; the fraction over the bound is the measure that
 * plain steps pass. */
int hd_z_check(hd_ctx* c, double bound, double* max_abs_dz, double* frac_over, double* wet_volume,
               double* wet_volume_snap);

/* A second context with the same case and a copy of the whole device state
 * (fields, masks, segment and run tables, counters): for what-if runs and
 * for tuning on a scratch copy.  src must be between whole steps. */
int hd_clone(hd_ctx* src, hd_ctx** out);
/* Copy the whole device state of src into dst, a context of the same case
 * (a clone); both between whole steps. */
int hd_copy_state(hd_ctx* dst, hd_ctx* src);
/* Launch-geometry autotuner: solver blocks per line (1, 2, 4) x block width
 * of the per-row kernels (64, 128, 256), each timed over `steps` replayed
 * steps from t on a clone of c; c keeps the fastest (its state is untouched;
 * every choice gives identical results).  Outputs are optional. */
int hd_autotune(hd_ctx* c, double t, const hd_step_opts* o, int steps, int* rb_mult, int* bw, double* ms_per_step);
/* Set those two knobs directly (rb_mult 1..16, bw 64 / 128 / 256). */
int hd_set_launch(hd_ctx* c, int rb_mult, int bw);

/* Batched ensemble: independent cases (different grids, series, roughness,
 * parameters) on one device, stepped together so that every kernel of a step
 * is ONE launch over all of them (grid.z indexes the case; the packed wet-run
 * tables of all cases feed one solver launch).  This has no counterpart in
 * the reference, which runs one case per process (rep.py:145-173); it serves
 * ensembles of many scenarios (e.g. 64 rivers of 256-1024 cells a side), far
 * more work per launch than one case.  Members are
 * ordinary contexts, created, uploaded and hd_init'ed as usual and read back
 * with hd_download; they must share the device and dT, be at the same time
 * and have taken only whole steps (no pending half-step).  A member may not
 * be stepped on its own while it belongs to an ensemble.  Results are
 * identical, bit for bit, to stepping each member alone. */
typedef struct hd_ensemble hd_ensemble;
int hd_ensemble_create(hd_ctx* const* members, int n, hd_ensemble** out);
int hd_ensemble_destroy(hd_ensemble* e);
const char* hd_ensemble_last_error(const hd_ensemble* e);
/* nsteps steps of every member from t (same options for all); counted adds
 * each member's wet-cell count after every step to that member's counter
 * (hd_wet_cell_steps). */
int hd_ensemble_run(hd_ensemble* e, double t, int nsteps, const hd_step_opts* o, int counted, double* t_out);
int hd_ensemble_profile(hd_ensemble* e, int enable);
int hd_ensemble_kernel_times(hd_ensemble* e, int n, const char** names, double* total_ms, int* launches);

/* Stand-alone kernels for unit parity. */
/* Batched Thomas solve (UVZSolver_multithread.cu:203-228 applied per system):
 * nsys systems, system s occupies rows off[s] .. off[s]+sn[s] of the device
 * arrays A,B,C,D (lower, diagonal, upper, rhs); x receives the solution.
 * All pointers are host pointers; the call copies in and out. */
int hd_thomas_batch(int device, int nsys, const int* off, const int* sn,
                    const double* A, const double* B, const double* C,
                    const double* D, double* x, int nrows);
/* The same systems through the run solvers' speculative parallel Thomas
 * (one wave per system, rows in LDS chunks of 510; the tridiag of
 * UVZSolver_multithread.cu:203-228 that k_adi_spec runs): the same bits as
 * hd_thomas_batch, for unit tests of the speculation and its repairs on
 * chosen systems. */
int hd_thomas_batch_spec(int device, int nsys, const int* off, const int* sn,
                         const double* A, const double* B, const double* C,
                         const double* D, double* x, int nrows);

/* Per-kernel entry points for unit parity.  Each runs the library's kernel(s)
 * for the reference kernel it is named after on the context's current state
 * (between whole steps) and waits for them; the step entry points above are
 * the product path.  After one that changes the mask (Onetime_init,
 * Find_Calculation_limits, Reset_states) the next step re-derives its active
 * tiles; call hd_prepare first when the segment tables should follow a mask
 * changed otherwise. */
/* Onetime_init, support_funcs.cu:20-58 (rep.py:126-130). */
int hd_k_Onetime_init(hd_ctx* c);
/* Find_Calculation_limits_Horizontal / _Vertical with the FS edge
 * interpolations, support_funcs.cu:189-335 (rep.py:135-136). */
int hd_k_Find_Calculation_limits(hd_ctx* c);
/* Htuongdoi, support_funcs.cu:338-354. */
int hd_k_Htuongdoi(hd_ctx* c);
/* preprocess_data, support_funcs.cu:433-542 (rep.py:139). */
int hd_k_preprocess_data(hd_ctx* c);
/* Update_Boundary_Value at time t (support_funcs.cu:544-668), or with
 * analytic != 0 the channel tide boundary_up/down/left/right
 * (support_funcs.cu:356-429), Engine.py:29-46. */
int hd_k_Update_Boundary_Value(hd_ctx* c, double t, int analytic);
/* The UZ / VZ half-step solve of every line: _calculate_preindex, _abcd,
 * _matrix_coeff, tridiagSolver and _extract_solution
 * (UVZSolver_multithread.cu:203-723, 818-943; Engine.py:160-178, 269-296);
 * shift 1 is the narrow-channel start index (Engine.py:152-158, 273-279). */
int hd_k_UZSolver(hd_ctx* c, int shift);
int hd_k_VZSolver(hd_ctx* c, int shift);
/* The tridiagonal rows alone (UZ/VZSolver_calculate_matrix_coeff,
 * UVZSolver_multithread.cu:469-593), written to the caller's arrays in the
 * reference's layout: line L (column j for UZ, M+3 lines; row i for VZ, N+3
 * lines), row r of the run starting at first
 * at L*(2N+1) + 2*first + r for UZ, L*(2M+1) + 2*first + r for VZ; SN gets
 * the system sizes, (max(N,M)+3) x segment_limit.  Positions outside the
 * systems' rows hold unspecified values.  Any pointer may be NULL. */
int hd_k_UZSolver_calculate_matrix_coeff(hd_ctx* c, int shift, double* AA, double* BB, double* CC, double* DD,
                                         int* SN);
int hd_k_VZSolver_calculate_matrix_coeff(hd_ctx* c, int shift, double* AA, double* BB, double* CC, double* DD,
                                         int* SN);
/* Normalize + update_buffer of t_u (isU) or t_v, support_funcs.cu:682-729. */
int hd_k_Normalize(hd_ctx* c, int isU);
/* solveV + update_margin_elem_V (UVZSolver_multithread.cu:727-768,
 * 1008-1071), solveU + update_margin_elem_U (:773-815, 947-1006). */
int hd_k_solveV(hd_ctx* c);
int hd_k_solveU(hd_ctx* c);
/* update_h_moi, Reset_states_horizontal (horizontal != 0) or _vertical,
 * update_uvz with the narrow-channel flags and Htuongdoi, in that order
 * (support_funcs.cu:60-186, 671-680, 338-354; Engine.py:199-214). */
int hd_k_Reset_states(hd_ctx* c, int horizontal, int kenhhepd, int kenhhepng);
/* Find_VTH and hesoK, Sediment_Transport.cu:508-568. */
int hd_k_Find_VTH_hesoK(hd_ctx* c);
/* Scan_FSj, tridiagSolver, FSj_extract_solution and Update_FS at time t
 * (Sediment_Transport.cu:205-338, 594-629; Engine.py:218-239); FSi the
 * same along rows (:62-202, 570-591; Engine.py:338-361). */
int hd_k_FSj_solve(hd_ctx* c, double t, double s_start, int ketdinh);
int hd_k_FSi_solve(hd_ctx* c, double t, double s_start, int ketdinh);
/* The concentration rows alone (_FSj / _FSi_calculate__mactrix_coeff) in the
 * reference's layout: FSj row of cell (i, j) at j*(N+3) + i, FSi at
 * i*(M+3) + j; positions outside the systems' rows hold unspecified values. */
int hd_k_Scan_FSj(hd_ctx* c, double t, double s_start, int ketdinh, double* AA, double* BB, double* CC,
                  double* DD);
int hd_k_Scan_FSi(hd_ctx* c, double t, double s_start, int ketdinh, double* AA, double* BB, double* CC,
                  double* DD);
/* Calculate_Qb (Sediment_Transport.cu:341-414) and BedLoad over rows
 * [sidx, eidx) (:416-506, 630-639). */
int hd_k_Calculate_Qb(hd_ctx* c, int qb_mode);
int hd_k_BedLoad(hd_ctx* c, int ketdinh, int sidx, int eidx);

/* hd_thomas_batch without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * pointer is DEVICE memory the caller owns (off, sn: nsys ints; A, B, C, D,
 * x: nrows doubles); the work is enqueued on `stream` (a hipStream_t, or
 * NULL for the null stream) and the call returns without waiting.  The
 * elimination factors live in a static device array of the library, so
 * nrows may be at most HD_THOMAS_NOALLOC_MAX_ROWS (HD_EINVAL beyond it).
 * The tables stay on the device, so a system outside rows [0, nrows) is
 * skipped rather than reported.  Calls that overlap on one device share that
 * array and race: serialise them (one stream, or events). */
#define HD_THOMAS_NOALLOC_MAX_ROWS (1 << 20)
int hd_thomas_batch_noalloc(int device, int nsys, const int* off, const int* sn, const double* A,
                            const double* B, const double* C, const double* D, double* x, int nrows,
                            void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HYDIST_H */
