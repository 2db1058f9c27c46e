/*
 * hd_internal.h -- device data layout shared by the libhydist kernels.
 *
 * HBM layout (one context = one case on one GPU):
 *  - every 2-D field of the reference (Argument_Pointers, constant.cuh:6-20)
 *    is its own (N+3) x (M+3) row-major FP64 array, flat index i*(M+3)+j, so
 *    uploads and downloads are the caller's numpy arrays byte for byte;
 *  - t_u and t_v each have a second buffer: the smoothing pass
 *    (Normalize/update_buffer, support_funcs.cu:682-729) writes the other
 *    buffer and the two swap roles, instead of a copy-back kernel;
 *  - the tridiagonal rows of a half-step are stored line-major like the
 *    reference's AA/BB/CC/DD (UVZSolver_multithread.cu:489-492): line L, row r
 *    at L * tw + r, tw = 2*len+2, with the elimination factors Ap/Bp beside
 *    them; per-segment continuity coefficients at the two line ends live in
 *    a small side table (SEG_*).
 * Kernels receive HdDev by value (kernel arguments), so a captured HIP graph
 * replays with exactly the pointers of the step it recorded.
 */
#ifndef HD_INTERNAL_H
#define HD_INTERNAL_H

#include <hip/hip_runtime.h>
#include "../../include/hydist.h"
#include "hd_math.h"

/* Members of the device view of one context.  G qualifies the pointee of
 * every device pointer: empty in HdDev (host code, kernel arguments), the
 * global address space in HdDevG (in the device compile), the view kernels
 * read from dv[blockIdx.z],
 * so that loads and stores through these pointers are global instructions
 * rather than flat ones (a pointer loaded from memory is otherwise generic).
 * Layouts are identical.  Fields: the reference's Argument_Pointers arrays;
 * hmax[4]; tmp_u / tmp_v, the smoothing ping-pong partners of t_u / t_v;
 * AA..Bp line-major rows and factors; seg_end / seg_sn per-run side tables;
 * runs_u / runs_v packed wet runs (line, first, last, 0) of the UZ / VZ
 * sweeps and run_cnt [0] / [1] their entry counts, [2] / [3] the chunks the
 * last UZ / VZ sweep repaired; flags [0] segment overflow, [1] boundary
 * range, [4] / [5] the chunks the running UZ / VZ sweep repairs; wet_count
 * [0] run accumulator, [1] one-shot count; tdev
 * the device time; tiles / ntiles the active 64-column row tiles; wetb the
 * wet bits of the segment finder, one 64-bit word per 64-cell tile row
 * ((N+3) x ceil((M+3)/64) words, bit b of word (i, w) for cell
 * (i, 64w + b): khouot == 0 and no drying flag); kh0b the same layout for
 * khouot == 0 before the wetting pass of the half-step; cand_cnt / cand the
 * wetting candidates (64 per line); dryflag one byte per cell. */
#define HD_DEV_MEMBERS(G) \
    int N, M, W, SL, total_time, series_len; \
    hd_params P; \
    G double *h, *hsnham, *VISCOIDX, *u, *v, *z, *t_u, *t_v, *t_z, *Htdu, *Htdv, *H_moi, *htaiz, *htaiz_bd; \
    G double *Kx1, *Ky1, *Tsxw, *Tsyw, *FS, *tFS, *VTH, *Kx, *Ky, *Fw, *Qbx, *Qby, *dH; \
    G int* khouot; \
    G double *ubt, *ubp, *vbt, *vbd, *hi; \
    G double *bc_up, *bc_down, *bc_left, *bc_right, *CC_u, *CC_d, *CC_l, *CC_r; \
    G int *bienQ, *boundary_type, *daui, *cuoii, *dauj, *cuoij, *moci, *mocj; \
    G double* hmax; \
    G double *tmp_u, *tmp_v; \
    G double *AA, *BB, *CC, *DD, *Ap, *Bp; \
    G double* seg_end; \
    G int* seg_sn; \
    G int4 *runs_u, *runs_v; \
    G int* run_cnt; \
    G int* flags; \
    G long long* wet_count; \
    G double* tdev; \
    G int* tiles; \
    G int* ntiles; \
    G unsigned long long* wetb; \
    G unsigned long long* kh0b; \
    G int *cand_cnt, *cand; \
    G unsigned char* dryflag;

struct HdDev {
    HD_DEV_MEMBERS()
};

#if defined(__HIP_DEVICE_COMPILE__)
#define HD_GLOBAL_AS __attribute__((address_space(1)))
#else
#define HD_GLOBAL_AS
#endif

struct HdDevG {
    HD_DEV_MEMBERS(HD_GLOBAL_AS)
};

static_assert(sizeof(HdDev) == sizeof(HdDevG), "device view layout");

/* the device view of batch member blockIdx.z */
__device__ __forceinline__ const HdDevG& dev_view(const HdDev* dv)
{
    return reinterpret_cast<const HdDevG*>(dv)[blockIdx.z];
}

/* The entries of a packed run table (cnt of them) that this workgroup of a
 * one-wave-per-run solver takes, the table being dealt to the first G
 * workgroups of the launch: start, end and stride of its loop.
 * Workgroups are dealt to the eight XCDs in turn (workgroup b to XCD b % 8)
 * and neighbouring table entries are neighbouring lines (each block of the
 * table scan reserves its lines' entries together), so XCD x takes the x-th
 * eighth of the table: lines that share cache lines are solved behind the
 * same L2 instead of fetching each line from HBM eight times. */
struct RunSlice {
    int start, end, step;
};
__device__ __forceinline__ RunSlice run_slice(int cnt, int G = (int)gridDim.x)
{
    const int nx = G < 8 ? 1 : 8;
    const int x = (int)blockIdx.x % nx, slot = (int)blockIdx.x / nx;
    const int per = (G - x + nx - 1) / nx; /* workgroups on XCD x */
    const int chunk = (cnt + nx - 1) / nx;
    const int lo = x * chunk;
    return RunSlice{lo + slot, cnt < lo + chunk ? cnt : lo + chunk, per};
}


/* A launch over a batch of contexts (one or an ensemble): dv points to n
 * device copies of HdDev, kernels take theirs from dv[blockIdx.z]; N, M are
 * the largest grid sides of the batch and size the launch. */
struct HdBatch {
    const HdDev* dv;
    int n, N, M;
};

/* kernel launchers (hd_kernels.hip, hd_sediment.hip), one kernel each */
void hd_launch_onetime_init(const HdBatch& b, hipStream_t s);
void hd_launch_find_limits(const HdBatch& b, hipStream_t s); /* wet bits from khouot, run-table reset, segments (+ FS edges) */
void hd_launch_uvz_segments(const HdBatch& b, int kd, int kn, int ntiles, int count, hipStream_t s);
void hd_launch_segments(const HdBatch& b, int count, hipStream_t s); /* also the FS edge interpolations */
void hd_launch_htuongdoi(const HdBatch& b, hipStream_t s);
void hd_launch_preprocess(const HdBatch& b, hipStream_t s);
void hd_launch_boundary_t(const HdBatch& b, int analytic, hipStream_t s);
void hd_launch_adi_coef(const HdBatch& b, int isU, int shift, hipStream_t s);
void hd_launch_adi_solve(const HdBatch& b, int isU, int shift, hipStream_t s);
void hd_launch_adi_line(const HdBatch& b, int isU, int shift, hipStream_t s);
void hd_launch_adi_run(const HdBatch& b, int isU, int shift, int nblocks, hipStream_t s);
/* bnd: the runs write the half-step's boundary values (series mode) */
void hd_launch_adi_spec(const HdBatch& b, int isU, int shift, int nblocks, int bnd, hipStream_t s);
void hd_launch_normalize_post(const HdBatch& b, int isU, int post, int horizontal, int bw, int ntiles, hipStream_t s);
void hd_launch_explicit(const HdBatch& b, int isU, hipStream_t s);
void hd_launch_margins(const HdBatch& b, int isU, hipStream_t s);
void hd_launch_explicit_margins(const HdBatch& b, int isU, int cells, hipStream_t s);
void hd_launch_explicit_rows_u(const HdBatch& b, int bw, int ntiles, hipStream_t s);
/* adv: advance the device time by dT/2 (after a sweep that wrote the boundary) */
void hd_launch_smooth_v_explicit_u(const HdBatch& b, int bw, int ntiles, int adv, hipStream_t s);
void hd_launch_smooth_u_explicit_v(const HdBatch& b, int bw, int ntiles, int adv, hipStream_t s);
/* mode: 0 the default wetting, 1 candidate windows for every line, 2 whole-line scans (cross-checks) */
void hd_launch_wetting(const HdBatch& b, int horizontal, int mode, hipStream_t s);
void hd_launch_smooth_wet(const HdBatch& b, int isU, int horizontal, int ntiles, int mode, hipStream_t s);
void hd_launch_uvz_dry(const HdBatch& b, int kd, int kn, int bw, int ntiles, hipStream_t s);
void hd_launch_vth_hesok(const HdBatch& b, int ntiles, hipStream_t s);
void hd_launch_fs_coef(const HdBatch& b, int which, double s_start, int ketdinh, hipStream_t s);
void hd_launch_fs_solve(const HdBatch& b, int which, hipStream_t s);
void hd_launch_fs_line(const HdBatch& b, int which, double s_start, int ketdinh, hipStream_t s);
void hd_launch_fs_run(const HdBatch& b, int which, double s_start, int ketdinh, int nblocks, int spec,
                      hipStream_t s);
void hd_launch_update_fs(const HdBatch& b, int ntiles, hipStream_t s);
void hd_launch_calc_qb(const HdBatch& b, int qb_mode, int ntiles, hipStream_t s);
void hd_launch_bedload(const HdBatch& b, int ketdinh, int slo, int eoff, int ntiles, hipStream_t s);
void hd_launch_wet_count(const HdBatch& b, hipStream_t s);
void hd_launch_flow_refresh(const HdBatch& b, hipStream_t s);
void hd_launch_bed_extrapolate(const HdBatch& b, const double* snap, double f, hipStream_t s);
void hd_launch_z_change(const HdBatch& b, const double* zref, double bound, double* out, hipStream_t s);
/* batched Thomas on device arrays; systems outside rows [0, nrows) are skipped */
int hd_thomas_batch_device(int nsys, const int* off, const int* sn, const double* A, const double* B,
                           const double* C, const double* D, double* x, double* Ap, double* Bp, int nrows,
                           hipStream_t s);
int hd_thomas_spec_device(int nsys, const int* off, const int* sn, const double* A, const double* B,
                           const double* C, const double* D, double* x, double* Ap, double* Bp, int nrows,
                           hipStream_t s);

#endif
