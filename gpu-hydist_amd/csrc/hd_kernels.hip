/*
 * hd_kernels.hip -- gfx950 kernels of the FP64 ADI step.
 *
 * One half-step (UZ shown; VZ is the transpose) is
 *   boundary -> coef (one thread per wet cell writes its two tridiagonal rows)
 *   -> solve (one lane per line: Thomas over each wet run, solution scattered
 *      straight into t_z / t_u during back substitution)
 *   -> smooth t_u -> explicit v -> v margins -> smooth t_v (+ H_moi, drying
 *      flags, wetting candidates) -> wetting (serial only at candidates)
 *   -> state copy (+ drying, face depths) -> segment tables (+ FS edge
 *      interpolation).
 * Every formula follows the reference kernel it replaces (cited per function)
 * in the same IEEE operation order; the code is built with -ffp-contract=off
 * so the bits match the CPU checker.  Wave size is 64.
 */
#include "hd_internal.h"
#include "hd_thomas.h"

#define W_ (d.W)

using namespace hd_spec;

namespace {

constexpr int kWave = 64;
constexpr int kCand = 64;   /* wetting candidates kept per line */

/* ================================================================ init */

/* Onetime_init, support_funcs.cu:20-58: every cell ends as khouot = 2 except
 * the cells of [1,N+1]x[1,M+1] whose 4-node mean depth exceeds NANGDAY. */
__global__ void k_onetime_init(const HdDev* __restrict__ dv)
{
    const HdDevG& d = dev_view(dv);
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    int N = d.N, M = d.M, W = d.W;
    if (i >= N + 3 || j >= M + 3) return;
    const double* h = d.h;
    int p = i * W + j;
    int kh = 2;
    if (i >= 1 && i <= N + 1 && j >= 1 && j <= M + 1) {
        double s4 = (h[(i - 1) * W + j - 1] + h[(i - 1) * W + j] + h[i * W + j - 1] + h[i * W + j]) * 0.25;
        if (s4 > d.P.NANGDAY) {
            kh = 0;
            d.H_moi[p] = 0;
        }
        if (i <= N && j <= M) {
            double ht = (h[(i - 1) * W + j - 1] + h[(i - 1) * W + j] + h[i * W + j - 1] + h[i * W + j]) * 0.25;
            d.htaiz[p] = ht;
            d.htaiz_bd[p] = ht;
            if (h[i * W + j - 1] + h[p] != 0)
                d.Kx1[p] = d.P.g * hd_pow((h[i * W + j - 1] + h[p]) * 0.5, -2 * d.P.mu_mn)
                           * hd_pow((d.hsnham[i * W + j - 1] + d.hsnham[p]) * 0.5, 2);
            if (h[(i - 1) * W + j] + h[p] != 0)
                d.Ky1[p] = d.P.g * hd_pow((h[(i - 1) * W + j] + h[p]) * 0.5, -2 * d.P.mu_mn)
                           * hd_pow((d.hsnham[(i - 1) * W + j] + d.hsnham[p]) * 0.5, 2);
        }
    }
    d.khouot[p] = kh;
}

/* Htuongdoi, support_funcs.cu:338-354 (init form, from z). */
__global__ void k_htuongdoi(const HdDev* __restrict__ dv)
{
    const HdDevG& d = dev_view(dv);
    int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int i = blockIdx.y + 1;
    int W = d.W;
    if (i > d.N || j > d.M) return;
    const double *h = d.h, *z = d.z;
    d.Htdu[i * W + j] = (h[i * W + j - 1] + h[i * W + j] + z[(i + 1) * W + j] + z[i * W + j]) * 0.5;
    d.Htdv[i * W + j] = (h[(i - 1) * W + j] + h[i * W + j] + z[i * W + j + 1] + z[i * W + j]) * 0.5;
}

/* One edge of preprocess_data, support_funcs.cu:450-511: hi = hbar^(2/3) /
 * (sum hbar^(5/3) dX) per wet run of the edge line, summed serially in index
 * order (the reference runs it single-threaded). */
__device__ void preprocess_edge(const HdDevG& d, int edge)
{
    int N = d.N, M = d.M, W = d.W, SL = d.SL;
    const double* h = d.h;
    double* hi = d.hi;
    const int *moc, *dau, *cuoi;
    int line;
    if (edge == 0) { moc = d.mocj; dau = d.dauj; cuoi = d.cuoij; line = M; }
    else if (edge == 1) { moc = d.mocj; dau = d.dauj; cuoi = d.cuoij; line = 2; hi += N + 3; }
    else if (edge == 2) { moc = d.moci; dau = d.daui; cuoi = d.cuoii; line = 2; hi += 2 * (N + 3); }
    else { moc = d.moci; dau = d.daui; cuoi = d.cuoii; line = N; hi += 2 * (N + 3) + (M + 3); }
    for (int k = 0; k < moc[line]; ++k) {
        double sum = 0;
        int a = dau[line * SL + k], b = cuoi[line * SL + k];
        for (int i = a; i <= b; ++i) {
            double v;
            if (edge == 0) v = (h[i * W + M] + h[(i - 1) * W + M]) / 2.0 - d.P.NANGDAY;
            else if (edge == 1) v = (h[i * W + 1] + h[(i - 1) * W + 1]) / 2.0 - d.P.NANGDAY;
            else if (edge == 2) v = (h[1 * W + i] + h[1 * W + i - 1]) / 2.0 - d.P.NANGDAY;
            else v = (h[N * W + i] + h[N * W + i - 1]) / 2.0 - d.P.NANGDAY;
            hi[i] = v;
            sum += hd_pow(v, 5.0 / 3.0);
        }
        for (int i = a; i <= b; ++i) hi[i] = hd_pow(hi[i], 2.0 / 3.0) / (sum * d.P.dX);
    }
}

/* preprocess_data, support_funcs.cu:433-542: the four edges (independent hi
 * ranges) in parallel, then hmax_* = 1/(ros * max htaiz_bd) per edge. */
__global__ void k_preprocess(const HdDev* __restrict__ dv)
{
    const HdDevG& d = dev_view(dv);
    int e = threadIdx.x;
    if (e < 4) preprocess_edge(d, e);
    if (e != 0) return;
    int N = d.N, M = d.M, W = d.W;
    const double* hb = d.htaiz_bd;
    double hu = hb[2 * W + M], hd = hb[2 * W + 2];
    for (int i = 3; i < N; ++i) {
        if (hb[i * W + M] > hu) hu = hb[i * W + M];
        if (hb[i * W + 2] > hd) hd = hb[i * W + 2];
    }
    double hl = hb[2 * W + 2], hr = hb[N * W + 2];
    for (int i = 3; i < M; ++i) {
        if (hb[2 * W + i] > hl) hl = hb[2 * W + i];
        if (hb[N * W + i] > hr) hr = hb[N * W + i];
    }
    if (hu != 0) hu = 1.0 / (d.P.ros * hu);
    if (hd != 0) hd = 1.0 / (d.P.ros * hd);
    if (hl != 0) hl = 1.0 / (d.P.ros * hl);
    if (hr != 0) hr = 1.0 / (d.P.ros * hr);
    d.hmax[0] = hu; d.hmax[1] = hd; d.hmax[2] = hl; d.hmax[3] = hr;
}

/* ================================================================ boundary */

/* boundary_up/down/left/right, support_funcs.cu:356-429 (analytic tide of the
 * narrow-channel mode), thread index i of one edge. */
__device__ void edge_analytic(const HdDevG& d, int edge, int i, double t)
{
    int N = d.N, M = d.M, W = d.W, SL = d.SL;
    const double PI = d.P.PI, dX = d.P.dX, dY = d.P.dY;
    if (edge == 0) {
        if (i < d.dauj[M * SL] || i > d.cuoij[M * SL]) return;
        d.t_z[i * W + M] = 0.01 * hd_cos(2 * PI / 27.750 * t) * hd_cos(2 * (PI / 100) * (100 - dY / 2));
        d.t_z[i * W + M + 1] = 0.01 * hd_cos(2 * PI / 27.75 * t) * hd_cos(2 * (PI / 100) * (100 + dY / 2));
    } else if (edge == 1) {
        if (i < d.dauj[2 * SL] || i > d.cuoij[2 * SL]) return;
        if (d.bienQ[1]) d.vbd[i] = 0;
        else {
            d.t_z[i * W + 2] = 0.01 * hd_cos(2 * PI / 27.75 * t) * hd_cos(2 * (PI / 100) * dY / 2);
            d.t_z[i * W + 1] = 0.01 * hd_cos(2 * PI / 27.75 * t) * hd_cos(2 * (PI / 100) * (-dY) / 2);
        }
    } else if (edge == 2) {
        if (i < d.daui[2 * SL] || i > d.cuoii[2 * SL]) return;
        if (d.bienQ[2]) d.ubt[i] = 0;
        else {
            d.t_z[2 * W + i] = 0.01 * hd_cos(2 * PI / 27.75 * t) * hd_cos(2 * (PI / 100) * dX / 2);
            d.t_z[1 * W + i] = 0.01 * hd_cos(2 * PI / 27.75 * t) * hd_cos(2 * (PI / 100) * (-dX) / 2);
        }
    } else {
        if (i < d.daui[2 * SL] || i > d.cuoii[N * SL]) return;
        if (d.bienQ[3]) d.ubp[i] = 0;
        else {
            d.t_z[N * W + i] = 0.01 * hd_cos(2 * PI / 27.75 * t) * hd_cos(2 * (PI / 100) * (100 - dX / 2));
            d.t_z[(N + 1) * W + i] = 0.01 * hd_cos(2 * PI / 27.75 * t) * hd_cos(2 * (PI / 100) * (100 + dX / 2));
        }
    }
}

/* Update_Boundary_Value, support_funcs.cu:622-668 (or the analytic channel
 * boundaries), one workgroup; the edges run in the reference's order
 * up, down, left, right, so a corner cell keeps the later edge's value.
 * The half-step time t = t_prev + dT/2 is advanced here, on the device.
 * Every thread reads the device time and its edge's tables in one round of
 * loads; the series values and the cells' weights follow in a second. */
__global__ void k_boundary(const HdDev* __restrict__ dv, int analytic)
{
    const HdDevG& d = dev_view(dv);
    double* tdev = d.tdev;
    const double t = tdev[0] + d.P.dT * 0.5;
    const int N = d.N, M = d.M, W = d.W, SL = d.SL, TT = d.total_time;
    const int imax = M > N ? M : N;
    if (analytic) {
        __syncthreads(); /* every thread has read the old time */
        if (threadIdx.x == 0) tdev[0] = t;
        int lo = 0, hi = imax + 2;
        for (int e = 0; e < 4; ++e) {
            for (int i = lo + threadIdx.x; i <= hi; i += blockDim.x) edge_analytic(d, e, i, t);
            __syncthreads();
        }
        return;
    }
    /* the four edges at once, 256 threads each, over the edge line's first
     * wet run (the only one the reference updates, support_funcs.cu:558) */
    const int e = threadIdx.x >> 8, lt = threadIdx.x & 255;
    const bool isU = e >= 2;
    const int* moc = isU ? d.moci : d.mocj;
    const int* dau = isU ? d.daui : d.dauj;
    const int* cuoi = isU ? d.cuoii : d.cuoij;
    const int location = e == 0 ? M : (e == 3 ? N : 2);
    const int loc_ext = e == 0 ? M + 1 : (e == 3 ? N + 1 : 1);
    const int mc = moc[location], a = dau[location * SL], b = cuoi[location * SL];
    const int bt0 = d.boundary_type[e * SL];
    const double hmax = d.hmax[e];
    /* row-edge coverage of the corner cells (rows 2 and N) of the column edges */
    const int m2 = d.moci[2], a2 = d.daui[2 * SL], b2 = d.cuoii[2 * SL], rbt2 = d.boundary_type[2 * SL];
    const int mN = d.moci[N], aN = d.daui[N * SL], bN = d.cuoii[N * SL], rbt3 = d.boundary_type[3 * SL];
    __syncthreads(); /* every thread has read the old time */
    if (threadIdx.x == 0) tdev[0] = t;
    const int t1 = (int)(t / 3600);
    const double t2 = (t - (3600.0 * t1)) / 3600.0;
    if (t1 < 0 || t1 + 1 >= TT || t1 + 1 >= d.series_len) {
        if (threadIdx.x == 0) d.flags[1] = 1;
        return;
    }
    if (mc <= 0) return;
    const double* bc = e == 0 ? d.bc_up : e == 1 ? d.bc_down : e == 2 ? d.bc_left : d.bc_right;
    const double* CCs = e == 0 ? d.CC_u : e == 1 ? d.CC_d : e == 2 ? d.CC_l : d.CC_r;
    const double* hi = d.hi + (e == 0 ? 0 : e == 1 ? (N + 3) : e == 2 ? 2 * (N + 3) : 2 * (N + 3) + (M + 3));
    double* barr = e == 0 ? d.vbt : e == 1 ? d.vbd : e == 2 ? d.ubt : d.ubp;
    for (int i = (a > 2 ? a : 2) + lt; i <= b && i <= imax; i += 256) {
        /* Boundary_value + FS_boundary, support_funcs.cu:544-617, segment 0 */
        const int fi = isU ? location : i, fj = isU ? i : location;
        const double b0 = bc[t1], b1 = bc[t1 + 1];
        const double c0 = CCs[t1], c1 = CCs[t1 + 1];
        const double hw = hi[i], hbd = d.htaiz_bd[fi * W + fj];
        /* a column edge (0, 1) leaves its cell of row 2 / N to the row edge
         * (2 / 3) that writes the same field there: in the reference's edge
         * order the later edge wins */
        bool yield = false;
        if (!isU && (i == 2 || i == N)) {
            const bool covers = (i == 2) ? (m2 > 0 && a2 <= location && location <= b2)
                                         : (mN > 0 && aN <= location && location <= bN);
            const int rb0 = (i == 2) ? rbt2 : rbt3;
            yield = covers && ((rb0 == 2) == (bt0 == 2));
        }
        const double bv = b0 * (1.0 - t2) + b1 * t2;
        if (bt0 == 2) {
            barr[i] = bv * hw;
        } else if (isU) {
            d.t_z[location * W + i] = bv;
            d.t_z[loc_ext * W + i] = bv;
        } else {
            if (!yield) d.t_z[i * W + location] = bv;
            d.t_z[i * W + loc_ext] = bv;
        }
        if (bt0 != 2 || yield) continue;
        const double cv = c0 * (1.0 - t2) + c1 * t2;
        d.FS[fi * W + fj] = cv * hbd * hmax;
    }
}

/* The device time step of k_boundary (t = tdev + dT/2, with its series
 * range flag) for a half-step whose boundary values the sweep's runs
 * wrote: one thread, in the launch after the sweep. */
__device__ __forceinline__ void advance_time(const HdDevG& d)
{
    const double t = d.tdev[0] + d.P.dT * 0.5;
    d.tdev[0] = t;
    const int t1 = (int)(t / 3600);
    if (t1 < 0 || t1 + 1 >= d.total_time || t1 + 1 >= d.series_len) d.flags[1] = 1;
}

/* The repaired-chunk count of the sweep just finished (direction 0 UZ, 1
 * VZ) becomes the one the next sweep of that direction reads; one thread,
 * in the launch after the sweep. */
__device__ __forceinline__ void roll_repairs(const HdDevG& d, int dir)
{
    d.run_cnt[2 + dir] = d.flags[4 + dir];
    d.flags[4 + dir] = 0;
}

/* The first wet run of each edge line (column M, column 2, row 2, row N):
 * its first cell a and its cells [lo, hi] clipped to [2, max(M, N)] as
 * k_boundary's loop (hi < lo when the line is dry). */
struct EdgeRuns {
    int lo[4], hi[4], a[4];
};

__device__ __forceinline__ EdgeRuns edge_runs(const HdDevG& d)
{
    const int N = d.N, M = d.M, SL = d.SL;
    const int imax = M > N ? M : N;
    EdgeRuns r;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const bool col = e < 2;
        const int loc = e == 0 ? M : (e == 3 ? N : 2);
        const int mc = col ? d.mocj[loc] : d.moci[loc];
        const int a = col ? d.dauj[loc * SL] : d.daui[loc * SL];
        const int b = col ? d.cuoij[loc * SL] : d.cuoii[loc * SL];
        r.a[e] = a;
        r.lo[e] = a > 2 ? a : 2;
        r.hi[e] = mc > 0 ? (b < imax ? b : imax) : r.lo[e] - 1;
    }
    return r;
}

/* One boundary value of the series mode (Boundary_value + FS_boundary,
 * support_funcs.cu:544-617, as k_boundary) of edge e (0 up: column M, 1
 * down: column 2, 2 left: row 2, 3 right: row N) at its index i (the row of
 * a column edge, the column of a row edge), at time t1 + t2 hours. */
__device__ __forceinline__ void edge_cell(const HdDevG& d, int e, int i, int t1, double t2)
{
    const int N = d.N, M = d.M, W = d.W, SL = d.SL;
    const bool col = e < 2;
    const int loc = e == 0 ? M : (e == 3 ? N : 2);
    const int ext = e == 0 ? M + 1 : (e == 3 ? N + 1 : 1);
    const int fi = col ? i : loc, fj = col ? loc : i;
    const double* bc = e == 0 ? d.bc_up : e == 1 ? d.bc_down : e == 2 ? d.bc_left : d.bc_right;
    const double* CCs = e == 0 ? d.CC_u : e == 1 ? d.CC_d : e == 2 ? d.CC_l : d.CC_r;
    const double* he = d.hi + (e == 0 ? 0 : e == 1 ? (N + 3) : e == 2 ? 2 * (N + 3) : 2 * (N + 3) + (M + 3));
    double* barr = e == 0 ? d.vbt : e == 1 ? d.vbd : e == 2 ? d.ubt : d.ubp;
    const int bt0 = d.boundary_type[e * SL];
    const double bv = bc[t1] * (1.0 - t2) + bc[t1 + 1] * t2;
    if (bt0 == 2) {
        barr[i] = bv * he[i];
        const double cv = CCs[t1] * (1.0 - t2) + CCs[t1 + 1] * t2;
        d.FS[fi * W + fj] = cv * d.htaiz_bd[fi * W + fj] * d.hmax[e];
    } else if (col) {
        d.t_z[i * W + loc] = bv;
        d.t_z[i * W + ext] = bv;
    } else {
        d.t_z[loc * W + i] = bv;
        d.t_z[ext * W + i] = bv;
    }
}

/* The half-step time of the boundary series: t = tdev + dT/2 as hours
 * t1 + t2; false when the series is too short (advance_time flags it). */
__device__ __forceinline__ bool boundary_time(const HdDevG& d, int& t1, double& t2)
{
    const double t = d.tdev[0] + d.P.dT * 0.5;
    t1 = (int)(t / 3600);
    t2 = (t - (3600.0 * t1)) / 3600.0;
    return !(t1 < 0 || t1 + 1 >= d.total_time || t1 + 1 >= d.series_len);
}

/* The boundary values of the series mode that one wet run of a sweep owns,
 * written by the run itself before it reads its end values, in place of a
 * k_boundary launch before the sweep.  Edge e's cells are those of the
 * first wet run of its edge line.  An edge along the sweep's lines (UZ:
 * columns M and 2; VZ: rows 2 and N) is owned by that line's first run,
 * which holds all of its cells.  A cell of an edge across the lines is
 * owned by the run of its line that starts (row or column 2) or ends (row N,
 * column M) there, when the line has one: the run rule ends a run at N or M
 * only when the two cells before are wet, so some edge cells lie in no run:
 * across_boundary writes those.  A run reads only its own
 * line's boundary values (t_z at its first / last cell and its line's
 * discharges, used only when the run starts at 2 or ends at N / M), so no
 * run reads a value that another run writes.  The edges are written in the
 * reference's order (0 up, 1 down, 2 left, 3 right) by the lane of the
 * cell, so at a corner the later edge's value stays, as k_boundary's yield
 * rule gives.  The device time is read here and advanced by the next launch
 * (k_smooth_*_explicit_*). */
template <bool isU>
__device__ __forceinline__ void run_boundary(const HdDevG& d, int line, int first, int last, int lane)
{
    const int N = d.N, M = d.M;
    /* only runs on an edge line or touching one can own a cell: most runs
     * leave here without a load */
    if (!(isU ? (line == M || line == 2 || first == 2 || last == N)
              : (line == 2 || line == N || first == 2 || last == M)))
        return;
    const EdgeRuns er = edge_runs(d);
    /* which cells k of this run each edge writes: [klo, khi] */
    int klo[4], khi[4];
    bool any = false, along = false;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const bool col = e < 2;
        const int loc = e == 0 ? M : (e == 3 ? N : 2);
        if (isU == col) {
            const bool mine = line == loc && first == er.a[e] && er.hi[e] >= er.lo[e];
            klo[e] = er.lo[e];
            khi[e] = mine ? er.hi[e] : er.lo[e] - 1;
            along = along || mine;
        } else {
            const bool mine = line >= er.lo[e] && line <= er.hi[e] && (loc == 2 ? first == 2 : last == loc);
            klo[e] = loc;
            khi[e] = mine ? loc : loc - 1;
        }
        any = any || khi[e] >= klo[e];
    }
    if (!any) return;
    int t1;
    double t2;
    if (!boundary_time(d, t1, t2)) return;
    auto cell = [&](int k) {
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (k >= klo[e] && k <= khi[e]) edge_cell(d, e, (isU == (e < 2)) ? k : line, t1, t2);
    };
    if (along) {
        for (int k = first + lane; k <= last; k += kWave) cell(k);
    } else if (lane == 0) {
        cell(first);
        if (last != first) cell(last);
    }
}

/* The cells of the two edges across a sweep's lines (UZ: rows 2 and N; VZ:
 * columns M and 2) that no run of the sweep holds, by one extra workgroup
 * of the sweep.  A cell is held when its line has a run that starts (at
 * row or column 2) or ends (at row N, column M) there: that run writes it
 * in run_boundary and may overwrite it with its solution afterwards, so it
 * is left alone here.  Runs are sorted along their line, so the line's
 * first table entry tells whether one starts at 2 and its last whether one
 * ends at N / M.  The other cells are read by no run of the sweep, only by
 * the launches after it, and no other edge writes them: a corner that a
 * row edge also covers lies in that row's first run, which then ends at M
 * or starts at 2 and holds it. */
template <bool isU>
__device__ __forceinline__ void across_boundary(const HdDevG& d, const EdgeRuns& er, int lane)
{
    const int N = d.N, M = d.M, SL = d.SL;
    int t1;
    double t2;
    if (!boundary_time(d, t1, t2)) return;
    /* the crossing lines' tables: columns for UZ, rows for VZ */
    const int* moc = isU ? d.mocj : d.moci;
    const int* dau = isU ? d.dauj : d.daui;
    const int* cuoi = isU ? d.cuoij : d.cuoii;
#pragma unroll
    for (int e = isU ? 2 : 0; e < (isU ? 4 : 2); ++e) {
        const int loc = e == 0 ? M : (e == 3 ? N : 2);
        for (int i = er.lo[e] + lane; i <= er.hi[e]; i += kWave) {
            const int mc = moc[i] < SL ? moc[i] : SL;
            const bool held = mc > 0 && (loc == 2 ? dau[i * SL] == 2 : cuoi[i * SL + mc - 1] == loc);
            if (!held) edge_cell(d, e, i, t1, t2);
        }
    }
}

/* ================================================================ ADI rows */

struct Pre {
    double f1, f2, f3, f5, a2, c2, d2;
};

/* Every value a preindex cell may read, loaded unconditionally so that all
 * loads of a cell (20, all in bounds for 2 <= i <= N, 2 <= j <= M) are in
 * flight together; the evaluation below takes the same branches and performs
 * the same operations as the reference on them. */
struct PreRaw {
    double c, n1, n2, m1, m2;   /* along-direction velocity at the cell and its +-1, +-2 cross neighbours */
    double x0, x1, x2, x3;      /* the four cross velocities averaged at the face */
    double vis, ht, ht_m, k1;   /* VISCOIDX, face depth here and one position back, Kx1 | Ky1 */
    double hm_m1, hm_p1, hm_m2; /* H_moi at -1, +1, -2 */
    double ts, zz, hc0, hc1;    /* Tsxw | Tsyw, z, cross face depths */
};

/* loads of _uzSolver_calculate_preindex (UVZSolver_multithread.cu:352-407) at (i, j) */
__device__ __forceinline__ PreRaw pre_uz_ld(const HdDevG& d, int i, int j)
{
    const int W = d.W, p = i * W + j;
    PreRaw r;
    r.c = d.u[p]; r.n1 = d.u[p + 1]; r.n2 = d.u[p + 2]; r.m1 = d.u[p - 1]; r.m2 = d.u[p - 2];
    r.x0 = d.v[p - 1]; r.x1 = d.v[p]; r.x2 = d.v[p + W - 1]; r.x3 = d.v[p + W];
    r.vis = d.VISCOIDX[p]; r.ht = d.Htdu[p]; r.ht_m = d.Htdu[p - W]; r.k1 = d.Kx1[p];
    r.hm_m1 = d.H_moi[p - 1]; r.hm_p1 = d.H_moi[p + 1]; r.hm_m2 = d.H_moi[p - 2];
    r.ts = d.Tsxw[p]; r.zz = d.z[p]; r.hc0 = d.Htdv[p]; r.hc1 = d.Htdv[p - 1];
    return r;
}

/* _uzSolver_calculate_preindex, UVZSolver_multithread.cu:352-407, on the
 * loaded values of cell (i, j). */
__device__ __forceinline__ Pre pre_uz_ev(const hd_params& P, const PreRaw& r)
{
    Pre o;
    double p = 0.0, q = 0.0;
    const double uc = r.c;
    const double vtb = (r.x0 + r.x1 + r.x2 + r.x3) * 0.25;
    const double vis = r.vis, htu = r.ht;
    o.f1 = P.dTchia2dX * uc + vis * P.dT / P.dXbp;
    o.f2 = -(2 + r.k1 * P.dT * sqrt(uc * uc + vtb * vtb) / htu + (2 * P.dT * vis) / P.dXbp);
    o.f3 = P.dT * vis / P.dXbp - P.dTchia2dX * uc;
    if (r.hm_m1 <= P.H_TINH) {
        if (vtb < 0) {
            p = vtb * (-3 * uc + 4 * r.n1 - r.n2) / P.dY2;
            q = (uc - 2 * r.n1 + r.n2) / P.dYbp;
        }
    } else {
        if (r.hm_p1 <= P.H_TINH) {
            if ((r.hm_m2 > P.H_TINH) && (vtb > 0)) {
                p = vtb * (3 * uc - 4 * r.m1 + r.m2) / P.dY2;
                q = (uc - 2 * r.m1 + r.m2) / P.dYbp;
            }
        } else {
            p = vtb * (r.n1 - r.m1) / P.dY2;
            q = (r.n1 - 2 * uc + r.m1) / P.dYbp;
        }
    }
    o.f5 = -2 * uc + P.dT * p - P.dT * P.CORIOLIS_FORCE * vtb - P.dT * vis * q - P.dT * (P.Windx - r.ts) / htu;
    o.c2 = P.dTchia2dX * htu;
    o.a2 = -P.dTchia2dX * r.ht_m;
    o.d2 = r.zz - P.dTchia2dY * (r.hc0 * r.x1 - r.hc1 * r.x0);
    return o;
}

/* loads of _vzSolver_calculate_preindex (UVZSolver_multithread.cu:295-350) at (i, j) */
__device__ __forceinline__ PreRaw pre_vz_ld(const HdDevG& d, int i, int j)
{
    const int W = d.W, p = i * W + j;
    PreRaw r;
    r.c = d.v[p]; r.n1 = d.v[p + W]; r.n2 = d.v[p + 2 * W]; r.m1 = d.v[p - W]; r.m2 = d.v[p - 2 * W];
    r.x0 = d.u[p - W]; r.x1 = d.u[p]; r.x2 = d.u[p - W + 1]; r.x3 = d.u[p + 1];
    r.vis = d.VISCOIDX[p]; r.ht = d.Htdv[p]; r.ht_m = d.Htdv[p - 1]; r.k1 = d.Ky1[p];
    r.hm_m1 = d.H_moi[p - W]; r.hm_p1 = d.H_moi[p + W]; r.hm_m2 = d.H_moi[p - 2 * W];
    r.ts = d.Tsyw[p]; r.zz = d.z[p]; r.hc0 = d.Htdu[p]; r.hc1 = d.Htdu[p - W];
    return r;
}

/* _vzSolver_calculate_preindex, UVZSolver_multithread.cu:295-350, on the
 * loaded values of cell (i, j). */
__device__ __forceinline__ Pre pre_vz_ev(const hd_params& P, const PreRaw& r)
{
    Pre o;
    double p = 0.0, q = 0.0;
    const double vc = r.c;
    const double utb = (r.x0 + r.x1 + r.x2 + r.x3) * 0.25;
    const double vis = r.vis, htv = r.ht;
    o.f1 = P.dTchia2dY * vc + vis * P.dT / P.dYbp;
    o.f2 = -(2 + r.k1 * P.dT * sqrt(vc * vc + utb * utb) / htv + (2 * P.dT * vis) / P.dYbp);
    o.f3 = P.dT * vis / P.dYbp - P.dTchia2dY * vc;
    if (r.hm_m1 <= P.H_TINH) {
        if (utb < 0) {
            q = utb * (-3 * vc + 4 * r.n1 - r.n2) / P.dX2;
            p = (vc - 2 * r.n1 + r.n2) / P.dXbp;
        }
    } else {
        if (r.hm_p1 <= P.H_TINH) {
            if ((r.hm_m2 > P.H_TINH) && (utb > 0)) {
                q = utb * (3 * vc - 4 * r.m1 + r.m2) / P.dX2;
                p = (vc - 2 * r.m1 + r.m2) / P.dXbp;
            }
        } else {
            q = utb * (r.n1 - r.m1) / P.dX2;
            p = (r.n1 - 2 * vc + r.m1) / P.dXbp;
        }
    }
    o.f5 = -2 * vc + P.dT * q + P.dT * P.CORIOLIS_FORCE * utb - P.dT * vis * p - P.dT * (P.Windy - r.ts) / htv;
    o.c2 = P.dTchia2dY * htv;
    o.a2 = -P.dTchia2dY * r.ht_m;
    o.d2 = r.zz - P.dTchia2dX * (r.hc0 * r.x1 - r.hc1 * r.x0);
    return o;
}

__device__ Pre pre_uz(const HdDevG& d, int i, int j) { return pre_uz_ev(d.P, pre_uz_ld(d, i, j)); }
__device__ Pre pre_vz(const HdDevG& d, int i, int j) { return pre_vz_ev(d.P, pre_vz_ld(d, i, j)); }

enum { S_WALL = 0, S_Q = 1, S_OPEN = 2 };
enum { E_Q = 1, E_OPEN = 2, E_WALL = 3 };

/* Start/end conditions of one run, _calculate_matrix_coeff
 * (UVZSolver_multithread.cu:469-593): start kind, end kind and system size
 * sn.  dkQ1 / dkQ2 are the Q flags of the start / end edges; at a non-wall
 * start the end test reads dkQ1, as UVZSolver_multithread.cu:562 is written. */
__device__ __forceinline__ void run_kinds(bool br1, bool br2, bool dkQ1, bool dkQ2, int first, int last, int dkfr,
                                          int& skind, int& ekind, int& sn)
{
    sn = 2 * (last - first);
    if (br1) {
        skind = S_WALL;
        if (!br2) {
            if (dkQ2 && last == dkfr) ekind = E_Q;
            else { sn--; ekind = E_OPEN; }
        } else ekind = E_WALL;
    } else {
        bool isBienran = dkQ1 && (first == 2);
        skind = isBienran ? S_Q : S_OPEN;
        if (!br2) {
            if (dkQ1 && last == dkfr) { if (!isBienran) sn--; ekind = E_Q; }
            else { sn--; if (!isBienran) sn -= 1; ekind = E_OPEN; }
        } else { if (!isBienran) sn--; ekind = E_WALL; }
    }
}

struct Mom {
    double a1, b1, c1, d1;
};

/* _calculate_abcd, UVZSolver_multithread.cu:409-466, at position k of the run
 * [first, last], given the preindex values of k and k+1.  The wall overrides
 * (first / last-1) replace the general formula. */
__device__ __forceinline__ Mom abcd(const Pre& a, const Pre& b, double f4, int k, int first, int last, bool br1, bool br2)
{
    Mom m;
    if (last - 1 > first) {
        if (br1 && k == first) {
            m.a1 = f4;
            m.b1 = a.f2 - (a.f3 * b.a2 / b.c2);
            m.c1 = -f4 - (a.f3 / b.c2);
            m.d1 = a.f5 - (a.f3 * b.d2 / b.c2);
        } else if (br2 && k == last - 1) {
            m.a1 = f4 - a.f1 / a.a2;
            m.b1 = a.f2 - (a.f1 * a.c2 / a.a2);
            m.c1 = -f4;
            m.d1 = a.f5 - (a.f1 * a.d2 / a.a2);
        } else {
            m.a1 = f4 - a.f1 / a.a2;
            m.c1 = -f4 - (a.f3 / b.c2);
            m.b1 = a.f2 - (a.f1 * a.c2 / a.a2) - (a.f3 * b.a2 / b.c2);
            m.d1 = a.f5 - (a.f1 * a.d2 / a.a2) - (a.f3 * b.d2 / b.c2);
        }
    } else {
        m.a1 = f4;
        m.b1 = a.f2;
        m.d1 = a.f5;
        m.c1 = -f4;
    }
    return m;
}

/* Locate the run of `line` holding position k in an LDS copy of the table. */
__device__ __forceinline__ int find_run(const int* dau, const int* cuoi, int n, int k, int& first, int& last)
{
    for (int s = 0; s < n; ++s) {
        int a = dau[s], b = cuoi[s];
        if (a <= k && k <= b) { first = a; last = b; return s; }
    }
    return -1;
}

/* Wall flags of locate_segment_u (column line) / locate_segment_v (row line),
 * constant.cuh:135-178. */
__device__ __forceinline__ void wall_flags(const HdDevG& d, bool isU, int line, int first, int last, bool& br1, bool& br2)
{
    int W = d.W;
    const double* h = d.h;
    double ND = d.P.NANGDAY;
    if (isU) {
        br1 = (first > 2) || ((first == 2) && ((h[1 * W + line] + h[1 * W + line - 1]) * 0.5 == ND));
        br2 = (last < d.N) || ((last == d.N) && ((h[d.N * W + line] + h[d.N * W + line - 1]) * 0.5 == ND));
    } else {
        br1 = (first > 2) || ((first == 2) && ((h[line * W + first - 1] + h[(line - 1) * W + first - 1]) * 0.5 == ND));
        br2 = (last < d.M) || ((last == d.M) && ((h[line * W + last] + h[(line - 1) * W + last]) * 0.5 == ND));
    }
}

/* Tridiagonal rows of one wet cell.  isU: UZ (line = column j, position = row
 * i) or VZ (line = row i, position = column j).  Rows of run [f, l] sit at
 * absolute index 2k (+1) of the line, as in the reference (offset = 2 first).
 * Layout: wall/Q start: row 2k = continuity of k, 2k+1 = momentum of k; open
 * start: 2k = momentum of k, 2k+1 = continuity of k+1. */
template <bool isU>
__global__ void __launch_bounds__(256) k_adi_coef(const HdDev* __restrict__ dv, int shift)
{
    const HdDevG& d = dev_view(dv);
    const int sidx = 2 + shift, eidx = isU ? d.M - shift : d.N;
    __shared__ int s_moc[64];
    __shared__ int s_dau[64][24];
    __shared__ int s_cuoi[64][24];
    int N = d.N, M = d.M, W = d.W, SL = d.SL;
    int tx = threadIdx.x, ty = threadIdx.y;
    int j0 = blockIdx.x * 64;
    int i0 = blockIdx.y * 16;
    /* lines held in LDS: columns j0..j0+63 (UZ) or rows i0..i0+15 (VZ) */
    int tid = ty * 64 + tx;
    if (isU) {
        for (int t = tid; t < 64; t += 256) {
            int L = j0 + t;
            s_moc[t] = (L >= sidx && L <= eidx && L <= M + 2) ? min(d.mocj[L], SL) : 0;
        }
        __syncthreads();
        for (int t = tid; t < 64 * SL; t += 256) {
            int c = t / SL, s = t % SL;
            if (s < s_moc[c]) { s_dau[c][s] = d.dauj[(j0 + c) * SL + s]; s_cuoi[c][s] = d.cuoij[(j0 + c) * SL + s]; }
        }
    } else {
        for (int t = tid; t < 16; t += 256) {
            int L = i0 + t;
            s_moc[t] = (L >= sidx && L <= eidx && L <= N + 2) ? min(d.moci[L], SL) : 0;
        }
        __syncthreads();
        for (int t = tid; t < 16 * SL; t += 256) {
            int c = t / SL, s = t % SL;
            if (s < s_moc[c]) { s_dau[c][s] = d.daui[(i0 + c) * SL + s]; s_cuoi[c][s] = d.cuoii[(i0 + c) * SL + s]; }
        }
    }
    __syncthreads();
    int j = j0 + tx;
    if (j > M + 2) return;
    const double f4 = isU ? 2 * d.P.g * d.P.dTchia2dX : 2 * d.P.g * d.P.dTchia2dY;
    const int tw = 2 * (N > M ? N : M) + 4;
    for (int r = ty; r < 16; r += 4) {
        int i = i0 + r;
        if (i > N + 2) break;
        int line = isU ? j : i;
        int k = isU ? i : j;
        int li = isU ? tx : r;
        int first, last;
        int s = find_run(s_dau[li], s_cuoi[li], s_moc[li], k, first, last);
        if (s < 0) continue;
        bool br1, br2;
        wall_flags(d, isU, line, first, last, br1, br2);
        bool dkQ1 = isU ? d.bienQ[2] != 0 : d.bienQ[1] != 0;
        bool dkQ2 = isU ? d.bienQ[3] != 0 : d.bienQ[0] != 0;
        int dkfr = isU ? N : M;
        int skind, ekind, sn;
        run_kinds(br1, br2, dkQ1, dkQ2, first, last, dkfr, skind, ekind, sn);
        Pre a = isU ? pre_uz(d, k, line) : pre_vz(d, line, k);
        Pre b = a;
        if (k + 1 <= last) b = isU ? pre_uz(d, k + 1, line) : pre_vz(d, line, k + 1);
        long long seg = (long long)line * SL + s;
        if (k == first) {
            d.seg_end[seg * 6 + 0] = a.a2;
            d.seg_end[seg * 6 + 1] = a.c2;
            d.seg_end[seg * 6 + 2] = a.d2;
            d.seg_sn[seg] = sn;
        }
        if (k == last) {
            d.seg_end[seg * 6 + 3] = a.a2;
            d.seg_end[seg * 6 + 4] = a.c2;
            d.seg_end[seg * 6 + 5] = a.d2;
        }
        if (sn < 0) continue;
        double ub1 = isU ? d.ubt[line] : d.vbd[line];
        double ub2 = isU ? d.ubp[line] : d.vbt[line];
        double TZ_l = isU ? d.t_z[first * W + line] : d.t_z[line * W + first];
        double TZ_r = isU ? d.t_z[last * W + line] : d.t_z[line * W + last];
        size_t base = (size_t)line * tw;
        for (int e = 0; e < 2; ++e) {
            int rr = 2 * (k - first) + e; /* system row */
            if (rr > sn) break;
            double A, B, C, D;
            bool cont = (skind != S_OPEN) ? (e == 0) : (e == 1);
            if (cont) {
                const Pre& c = (skind != S_OPEN) ? a : b;
                A = c.a2; B = 1; C = c.c2; D = c.d2;
            } else {
                Mom m = abcd(a, b, f4, k, first, last, br1, br2);
                A = m.a1; B = m.b1; C = m.c1; D = m.d1;
            }
            if (rr == 0) {
                if (skind == S_Q) D = a.d2 - a.a2 * ub1;
                else if (skind == S_OPEN) D = D - A * TZ_l;
            }
            if (rr == sn) {
                if (ekind == E_Q) { /* row is the continuity row of `last` */
                    const Pre& c = (k == last) ? a : b;
                    A = c.a2; B = 1; D = c.d2 - c.c2 * ub2;
                } else if (ekind == E_OPEN) { /* momentum row of last-1 (this cell) */
                    Mom m = abcd(a, b, f4, k, first, last, br1, br2);
                    A = m.a1; B = m.b1; D = m.d1 - m.c1 * TZ_r;
                } else {
                    const Pre& c = (k == last) ? a : b;
                    A = c.a2; B = 1; D = c.d2;
                }
            }
            size_t o = base + 2 * k + e;
            d.AA[o] = A; d.BB[o] = B; d.CC[o] = C; d.DD[o] = D;
        }
    }
}

/* tridiagSolver / tridiag (UVZSolver_multithread.cu:203-262) and the
 * extraction of _uz/_vzSolver_extract_solution (:595-723), one lane per line:
 * forward elimination over each run, back substitution writing t_z and t_u
 * (UZ) or t_v (VZ) as each unknown is resolved. */
template <bool isU>
__global__ void __launch_bounds__(64) k_adi_solve(const HdDev* __restrict__ dv, int shift)
{
    const HdDevG& d = dev_view(dv);
    const int sidx = 2 + shift, eidx = isU ? d.M - shift : d.N;
    int line = sidx + blockIdx.x * blockDim.x + threadIdx.x;
    if (line > eidx) return;
    int N = d.N, M = d.M, W = d.W, SL = d.SL;
    const int tw = 2 * (N > M ? N : M) + 4;
    int nseg = isU ? d.mocj[line] : d.moci[line];
    if (nseg > SL) nseg = SL;
    const int* dau = (isU ? d.dauj : d.daui) + line * SL;
    const int* cuoi = (isU ? d.cuoij : d.cuoii) + line * SL;
    bool dkQ1 = isU ? d.bienQ[2] != 0 : d.bienQ[1] != 0;
    bool dkQ2 = isU ? d.bienQ[3] != 0 : d.bienQ[0] != 0;
    bool exQ = isU ? d.bienQ[3] != 0 : d.bienQ[0] != 0;  /* extraction end test: bienQ[3] / bienQ[0] */
    bool exQs = isU ? d.bienQ[2] != 0 : d.bienQ[1] != 0; /* extraction start test: bienQ[2] / bienQ[1] */
    int dkfr = isU ? N : M;
    double ub1 = isU ? d.ubt[line] : d.vbd[line];
    double ub2 = isU ? d.ubp[line] : d.vbt[line];
    double* tz = d.t_z;
    double* tq = isU ? d.t_u : d.t_v;
    /* element (position k) of this line */
    auto at = [&](int k) -> size_t { return isU ? (size_t)k * W + line : (size_t)line * W + k; };
    size_t base = (size_t)line * tw;
    const double *AA = d.AA + base, *BB = d.BB + base, *CC = d.CC + base, *DD = d.DD + base;
    double *Ap = d.Ap + base, *Bp = d.Bp + base;
    for (int s = 0; s < nseg; ++s) {
        int first = dau[s], last = cuoi[s];
        bool br1, br2;
        wall_flags(d, isU, line, first, last, br1, br2);
        int skind, ekind, sn;
        run_kinds(br1, br2, dkQ1, dkQ2, first, last, dkfr, skind, ekind, sn);
        (void)ekind;
        const double* se = d.seg_end + ((size_t)line * SL + s) * 6;
        double q_last_prev = 0; /* t_u/t_v at last-1 after extraction */
        bool have_prev = false;
        double zl_sol = 0;
        if (sn >= 0) {
            int o = 2 * first;
            double x;
            if (sn == 0) {
                x = DD[o] / BB[o];
            } else {
                double ap = -CC[o] / BB[o];
                double bp = DD[o] / BB[o];
                Ap[o] = ap; Bp[o] = bp;
                for (int r = 1; r < sn; ++r) {
                    double a = AA[o + r], b = BB[o + r], c = CC[o + r], dd = DD[o + r];
                    double e = a * ap + b;
                    double nap = -c / e;
                    double nbp = (dd - (a * bp)) / e;
                    ap = nap; bp = nbp;
                    Ap[o + r] = ap; Bp[o + r] = bp;
                }
                double a = AA[o + sn];
                x = (DD[o + sn] - (a * bp)) / (BB[o + sn] + (a * ap));
            }
            /* back substitution + extraction, rows sn .. 0 */
            for (int r = sn; r >= 0; --r) {
                if (r < sn) x = Bp[o + r] + (Ap[o + r] * x);
                int k;
                bool isz;
                if (skind != S_OPEN) { k = first + (r >> 1); isz = (r & 1) == 0; }
                else { isz = (r & 1) == 1; k = first + ((r + 1) >> 1); }
                if (k < last) {
                    if (isz) tz[at(k)] = x;
                    else {
                        tq[at(k)] = x;
                        if (k == last - 1) { q_last_prev = x; have_prev = true; }
                    }
                } else if (isz && r == sn) {
                    zl_sol = x;
                }
            }
        }
        /* start of the run */
        if (first < last) {
            if (br1) tq[at(first - 1)] = 0;
            else if (exQs && first == 2) tq[at(first - 1)] = ub1;
            else if (sn >= 0) {
                double a2f = se[0], c2f = se[1], d2f = se[2];
                tq[at(first - 1)] = (d2f - tz[at(first)] - c2f * tq[at(first)]) / a2f;
            }
        }
        /* end of the run */
        if (br2) {
            tq[at(last)] = 0;
            if (sn >= 0) tz[at(last)] = zl_sol;
        } else if (exQ && last == dkfr) {
            tq[at(last)] = ub2;
            if (sn >= 0) tz[at(last)] = zl_sol;
        } else {
            double prev = have_prev ? q_last_prev : tq[at(last - 1)];
            double a2l = se[3], c2l = se[4], d2l = se[5];
            tq[at(last)] = (d2l - a2l * prev - tz[at(last)]) / c2l;
        }
    }
}

/* The two tridiagonal rows (e = 0, 1) of position k of run [first, last],
 * in the reference's interleaving (see k_adi_coef), with the start / end
 * conditions of _calculate_matrix_coeff applied (end row wins). */
struct Row4 {
    double A, B, C, D;
};

__device__ __forceinline__ Row4 cell_row(const Pre& a, const Pre& b, double f4, int k, int e, int first, int last,
                                         bool br1, bool br2, int skind, int ekind, int sn, double ub1, double ub2,
                                         double TZ_l, double TZ_r)
{
    int rr = 2 * (k - first) + e;
    Row4 w;
    bool cont = (skind != S_OPEN) ? (e == 0) : (e == 1);
    if (cont) {
        const Pre& c = (skind != S_OPEN) ? a : b;
        w.A = c.a2; w.B = 1; w.C = c.c2; w.D = c.d2;
    } else {
        Mom m = abcd(a, b, f4, k, first, last, br1, br2);
        w.A = m.a1; w.B = m.b1; w.C = m.c1; w.D = m.d1;
    }
    if (rr == 0) {
        if (skind == S_Q) w.D = a.d2 - a.a2 * ub1;
        else if (skind == S_OPEN) w.D = w.D - w.A * TZ_l;
    }
    if (rr == sn) {
        if (ekind == E_Q) {
            const Pre& c = (k == last) ? a : b;
            w.A = c.a2; w.B = 1; w.D = c.d2 - c.c2 * ub2;
        } else if (ekind == E_OPEN) {
            Mom m = abcd(a, b, f4, k, first, last, br1, br2);
            w.A = m.a1; w.B = m.b1; w.D = m.d1 - m.c1 * TZ_r;
        } else {
            const Pre& c = (k == last) ? a : b;
            w.A = c.a2; w.B = 1; w.D = c.d2;
        }
    }
    return w;
}

/* Fused UZ / VZ line solve, one wave per line (the reference's
 * _calculate_preindex + _calculate_abcd + _calculate_matrix_coeff +
 * tridiagSolver + _extract_solution, UVZSolver_multithread.cu:203-723).
 * For each wet run, chunks of 63 cells are assembled in parallel (lane k
 * computes the preindex of its cell into LDS, then its two rows from its own
 * and its neighbour's values), lane 0 carries the Thomas forward elimination
 * through the chunk's rows in LDS, then back-substitutes, and all lanes
 * scatter the solution into t_z / t_u (t_v).  Same operations in the same
 * order as the sequential reference Thomas, so the results are bit-identical;
 * only the schedule changes.  Elimination factors of rows beyond kCap spill
 * to the global Ap/Bp line buffers. */
template <bool isU>
__global__ void __launch_bounds__(64) k_adi_line(const HdDev* __restrict__ dv, int shift)
{
    const HdDevG& d = dev_view(dv);
    const int sidx = 2 + shift, eidx = isU ? d.M - shift : d.N;
    constexpr int CH = 63;
    constexpr int kCap = 2048;
    __shared__ double sP[7][64];
    __shared__ double sA[2 * CH], sB[2 * CH], sC[2 * CH], sD[2 * CH];
    __shared__ double sAp[kCap], sBp[kCap];
    __shared__ double s_end[6];
    int line = sidx + blockIdx.x;
    if (line > eidx) return;
    int lane = threadIdx.x;
    int N = d.N, M = d.M, W = d.W, SL = d.SL;
    const int tw = 2 * (N > M ? N : M) + 4;
    int nseg = isU ? d.mocj[line] : d.moci[line];
    if (nseg > SL) nseg = SL;
    const int* dau = (isU ? d.dauj : d.daui) + line * SL;
    const int* cuoi = (isU ? d.cuoij : d.cuoii) + line * SL;
    bool dkQ1 = isU ? d.bienQ[2] != 0 : d.bienQ[1] != 0;
    bool dkQ2 = isU ? d.bienQ[3] != 0 : d.bienQ[0] != 0;
    int dkfr = isU ? N : M;
    double ub1 = isU ? d.ubt[line] : d.vbd[line];
    double ub2 = isU ? d.ubp[line] : d.vbt[line];
    const double f4 = isU ? 2 * d.P.g * d.P.dTchia2dX : 2 * d.P.g * d.P.dTchia2dY;
    double* tz = d.t_z;
    double* tq = isU ? d.t_u : d.t_v;
    double* gAp = d.Ap + (size_t)line * tw;
    double* gBp = d.Bp + (size_t)line * tw;
    for (int s = 0; s < nseg; ++s) {
        int first = dau[s], last = cuoi[s];
        bool br1, br2;
        wall_flags(d, isU, line, first, last, br1, br2);
        int skind, ekind, sn;
        run_kinds(br1, br2, dkQ1, dkQ2, first, last, dkfr, skind, ekind, sn);
        double TZ_l = isU ? tz[(size_t)first * W + line] : tz[(size_t)line * W + first];
        double TZ_r = isU ? tz[(size_t)last * W + line] : tz[(size_t)line * W + last];
        double* oAp = gAp + 2 * first;
        double* oBp = gBp + 2 * first;
        double xs = 0;
        if (sn >= 0) {
            double ap = 0, bp = 0;
            for (int c0 = first; c0 <= last; c0 += CH) {
                int k = c0 + lane;
                Pre a;
                if (k <= last) {
                    a = isU ? pre_uz(d, k, line) : pre_vz(d, line, k);
                    sP[0][lane] = a.f1; sP[1][lane] = a.f2; sP[2][lane] = a.f3; sP[3][lane] = a.f5;
                    sP[4][lane] = a.a2; sP[5][lane] = a.c2; sP[6][lane] = a.d2;
                    if (k == first) { s_end[0] = a.a2; s_end[1] = a.c2; s_end[2] = a.d2; }
                    if (k == last) { s_end[3] = a.a2; s_end[4] = a.c2; s_end[5] = a.d2; }
                }
                __syncthreads();
                if (lane < CH && k <= last) {
                    Pre b = a;
                    if (k + 1 <= last) {
                        b.f1 = sP[0][lane + 1]; b.f2 = sP[1][lane + 1]; b.f3 = sP[2][lane + 1]; b.f5 = sP[3][lane + 1];
                        b.a2 = sP[4][lane + 1]; b.c2 = sP[5][lane + 1]; b.d2 = sP[6][lane + 1];
                    }
                    for (int e = 0; e < 2; ++e) {
                        int rr = 2 * (k - first) + e;
                        if (rr > sn) break;
                        Row4 w = cell_row(a, b, f4, k, e, first, last, br1, br2, skind, ekind, sn, ub1, ub2, TZ_l, TZ_r);
                        sA[2 * lane + e] = w.A; sB[2 * lane + e] = w.B; sC[2 * lane + e] = w.C; sD[2 * lane + e] = w.D;
                    }
                }
                __syncthreads();
                if (lane == 0) {
                    int rb = 2 * (c0 - first);
                    int re = rb + 2 * CH - 1;
                    if (re > sn) re = sn;
                    for (int r = rb; r <= re; ++r) {
                        double A = sA[r - rb], B = sB[r - rb], C = sC[r - rb], D = sD[r - rb];
                        if (r == 0 && sn == 0) {
                            xs = D / B;
                        } else if (r == 0) {
                            ap = -C / B;
                            bp = D / B;
                        } else if (r < sn) {
                            double e = A * ap + B;
                            double nap = -C / e;
                            double nbp = (D - (A * bp)) / e;
                            ap = nap;
                            bp = nbp;
                        } else {
                            xs = (D - (A * bp)) / (B + (A * ap));
                        }
                        if (r < sn) {
                            if (r < kCap) { sAp[r] = ap; sBp[r] = bp; }
                            else { oAp[r] = ap; oBp[r] = bp; }
                        }
                    }
                }
                __syncthreads();
            }
            /* back substitution; x[r] replaces Bp[r] */
            if (lane == 0) {
                double x = xs;
                if (sn < kCap) sBp[sn] = x; else oBp[sn] = x;
                for (int r = sn - 1; r >= 0; --r) {
                    double apr = r < kCap ? sAp[r] : oAp[r];
                    double bpr = r < kCap ? sBp[r] : oBp[r];
                    x = bpr + (apr * x);
                    if (r < kCap) sBp[r] = x; else oBp[r] = x;
                }
            }
            __syncthreads();
            /* scatter the unknowns of positions first .. last-1 */
            for (int r = lane; r <= sn; r += 64) {
                double x = r < kCap ? sBp[r] : oBp[r];
                int k;
                bool isz;
                if (skind != S_OPEN) { k = first + (r >> 1); isz = (r & 1) == 0; }
                else { isz = (r & 1) == 1; k = first + ((r + 1) >> 1); }
                if (k < last) {
                    size_t p = isU ? (size_t)k * W + line : (size_t)line * W + k;
                    if (isz) tz[p] = x; else tq[p] = x;
                }
            }
        } else if (lane == 0) {
            /* D21: no system; the cell's preindex is still needed at the end */
            Pre a = isU ? pre_uz(d, last, line) : pre_vz(d, line, last);
            s_end[3] = a.a2; s_end[4] = a.c2; s_end[5] = a.d2;
        }
        __syncthreads();
        if (lane == 0) {
            auto at = [&](int k) -> size_t { return isU ? (size_t)k * W + line : (size_t)line * W + k; };
            auto X = [&](int r) -> double { return r < kCap ? sBp[r] : oBp[r]; };
            bool exQ = dkQ2, exQs = dkQ1;
            if (first < last) {
                if (br1) tq[at(first - 1)] = 0;
                else if (exQs && first == 2) tq[at(first - 1)] = ub1;
                else if (sn >= 0) tq[at(first - 1)] = (s_end[2] - tz[at(first)] - s_end[1] * X(0)) / s_end[0];
            }
            if (br2) {
                tq[at(last)] = 0;
                if (sn >= 0) tz[at(last)] = X(sn);
            } else if (exQ && last == dkfr) {
                tq[at(last)] = ub2;
                if (sn >= 0) tz[at(last)] = X(sn);
            } else {
                double prev;
                if (first < last && sn >= 0) prev = X(skind != S_OPEN ? 2 * (last - 1 - first) + 1 : 2 * (last - 1 - first));
                else prev = tq[at(last - 1)];
                tq[at(last)] = (s_end[5] - s_end[3] * prev - tz[at(last)]) / s_end[4];
            }
        }
        __syncthreads();
    }
}

/* Fused UZ / VZ solve, one wave per wet run taken from the packed run table
 * that k_find_limits appends to (runs_u: column lines of the UZ half-step,
 * runs_v: row lines of the VZ half-step).  Same operations in the same order
 * as k_adi_line (and the reference's _calculate_preindex ... tridiagSolver ...
 * _extract_solution, UVZSolver_multithread.cu:203-723); runs of one line are
 * independent systems, so they now proceed on different waves and the
 * critical path of the launch is the longest run, not the longest line.
 * Chunks of 127 cells (254 rows) are assembled by the 64 lanes (two cells
 * each) into LDS; lane 0 carries the forward elimination through the chunk,
 * overwriting A/B with the factors ap/bp; the factors of all but the last
 * chunk go to the run's slice of the Ap/Bp line buffers and come back for
 * back substitution.  The end conditions take X(0), X(sn) and the last
 * interior u unknown from the solution held in LDS, and the unchanged boundary
 * depths TZ_l / TZ_r loaded at the start, so no lane reads a value another
 * lane of this launch wrote. */
template <bool isU>
__global__ void __launch_bounds__(64) k_adi_run(const HdDev* __restrict__ dv, int shift)
{
    const HdDevG& d = dev_view(dv);
    const int sidx = 2 + shift, eidx = isU ? d.M - shift : d.N;
    constexpr int CH = 127;
    constexpr int RW = 2 * CH;
    __shared__ double sP[7][CH + 1];
    __shared__ double sA[RW], sB[RW], sC[RW], sD[RW];
    __shared__ double sE[6];
    __shared__ double sX[3]; /* X(0), X(sn), X(last interior u row) */
    const int lane = threadIdx.x;
    const int N = d.N, M = d.M, W = d.W;
    const int tw = 2 * (N > M ? N : M) + 4;
    const int cap = (isU ? M + 3 : N + 3) * d.SL;
    const int cnt = d.run_cnt[isU ? 0 : 1] < cap ? d.run_cnt[isU ? 0 : 1] : cap;
    const int4* runs = isU ? d.runs_u : d.runs_v;
    const bool dkQ1 = isU ? d.bienQ[2] != 0 : d.bienQ[1] != 0;
    const bool dkQ2 = isU ? d.bienQ[3] != 0 : d.bienQ[0] != 0;
    const int dkfr = isU ? N : M;
    const double f4 = isU ? 2 * d.P.g * d.P.dTchia2dX : 2 * d.P.g * d.P.dTchia2dY;
    double* tz = d.t_z;
    double* tq = isU ? d.t_u : d.t_v;
    const RunSlice sl = run_slice(cnt);
    for (int ri = sl.start; ri < sl.end; ri += sl.step) {
        const int4 rn = runs[ri];
        const int line = rn.x, first = rn.y, last = rn.z;
        if (line < sidx || line > eidx) continue;
        auto at = [&](int k) -> size_t { return isU ? (size_t)k * W + line : (size_t)line * W + k; };
        bool br1, br2;
        wall_flags(d, isU, line, first, last, br1, br2);
        int skind, ekind, sn;
        run_kinds(br1, br2, dkQ1, dkQ2, first, last, dkfr, skind, ekind, sn);
        const double ub1 = isU ? d.ubt[line] : d.vbd[line];
        const double ub2 = isU ? d.ubp[line] : d.vbt[line];
        const double TZ_l = tz[at(first)];
        const double TZ_r = tz[at(last)];
        /* row of the u unknown of cell last-1 (read by the open-end rule) */
        const int ru = (skind != S_OPEN) ? 2 * (last - 1 - first) + 1 : 2 * (last - 1 - first);
        double* gAp = d.Ap + (size_t)line * tw + 2 * first;
        double* gBp = d.Bp + (size_t)line * tw + 2 * first;
        if (sn >= 0) {
            const int nch = sn / RW + 1; /* chunks of RW rows; cells of chunk c: first + CH*c ... */
            double ap = 0, bp = 0, xs = 0;
            for (int c = 0; c < nch; ++c) {
                const int c0 = first + CH * c;
                const int rb = RW * c;
                const int re = (rb + RW - 1 < sn) ? rb + RW - 1 : sn;
                Pre pa[2];
                for (int h = 0; h < 2; ++h) {
                    int t = lane + 64 * h;
                    int k = c0 + t;
                    if (k <= last) {
                        Pre a = isU ? pre_uz(d, k, line) : pre_vz(d, line, k);
                        pa[h] = a;
                        sP[0][t] = a.f1; sP[1][t] = a.f2; sP[2][t] = a.f3; sP[3][t] = a.f5;
                        sP[4][t] = a.a2; sP[5][t] = a.c2; sP[6][t] = a.d2;
                        if (k == first) { sE[0] = a.a2; sE[1] = a.c2; sE[2] = a.d2; }
                        if (k == last) { sE[3] = a.a2; sE[4] = a.c2; sE[5] = a.d2; }
                    }
                }
                __syncthreads();
                for (int h = 0; h < 2; ++h) {
                    int t = lane + 64 * h;
                    int k = c0 + t;
                    if (t < CH && k <= last) {
                        const Pre& a = pa[h];
                        Pre b = a;
                        if (k + 1 <= last) {
                            b.f1 = sP[0][t + 1]; b.f2 = sP[1][t + 1]; b.f3 = sP[2][t + 1]; b.f5 = sP[3][t + 1];
                            b.a2 = sP[4][t + 1]; b.c2 = sP[5][t + 1]; b.d2 = sP[6][t + 1];
                        }
                        for (int e = 0; e < 2; ++e) {
                            int rr = 2 * (k - first) + e;
                            if (rr > sn) break;
                            Row4 w = cell_row(a, b, f4, k, e, first, last, br1, br2, skind, ekind, sn, ub1, ub2, TZ_l,
                                              TZ_r);
                            int ix = rr - rb;
                            sA[ix] = w.A; sB[ix] = w.B; sC[ix] = w.C; sD[ix] = w.D;
                        }
                    }
                }
                __syncthreads();
                if (lane == 0) {
                    int i = 0;
                    if (rb == 0) {
                        if (sn == 0) {
                            xs = sD[0] / sB[0];
                        } else {
                            ap = -sC[0] / sB[0];
                            bp = sD[0] / sB[0];
                            sA[0] = ap; sB[0] = bp;
                        }
                        i = 1;
                    }
                    const int imid = ((re < sn - 1) ? re : sn - 1) - rb;
#pragma unroll 4
                    for (; i <= imid; ++i) {
                        double A = sA[i], B = sB[i], C = sC[i], D = sD[i];
                        double e = A * ap + B;
                        double nap = -C / e;
                        double nbp = (D - (A * bp)) / e;
                        ap = nap;
                        bp = nbp;
                        sA[i] = ap; sB[i] = bp;
                    }
                    if (re == sn && sn > 0) {
                        double A = sA[sn - rb], B = sB[sn - rb], D = sD[sn - rb];
                        xs = (D - (A * bp)) / (B + (A * ap));
                    }
                }
                __syncthreads();
                if (c + 1 < nch) { /* factors of this chunk (rows rb .. re, all < sn) to the line buffers */
                    for (int i = lane; i <= re - rb; i += 64) { gAp[rb + i] = sA[i]; gBp[rb + i] = sB[i]; }
                }
            }
            /* back substitution, last chunk first; x[r] replaces D[r] */
            for (int c = nch - 1; c >= 0; --c) {
                const int rb = RW * c;
                const int re = (rb + RW - 1 < sn) ? rb + RW - 1 : sn;
                if (c + 1 < nch) {
                    __syncthreads();
                    for (int i = lane; i <= re - rb; i += 64) { sA[i] = gAp[rb + i]; sB[i] = gBp[rb + i]; }
                    __syncthreads();
                }
                if (lane == 0) {
                    int i = re - rb;
                    if (c == nch - 1) { sD[i] = xs; --i; }
#pragma unroll 4
                    for (; i >= 0; --i) {
                        xs = sB[i] + (sA[i] * xs);
                        sD[i] = xs;
                    }
                }
                __syncthreads();
                /* scatter the unknowns of positions first .. last-1 */
                for (int i = lane; i <= re - rb; i += 64) {
                    int r = rb + i;
                    double x = sD[i];
                    int k;
                    bool isz;
                    if (skind != S_OPEN) { k = first + (r >> 1); isz = (r & 1) == 0; }
                    else { isz = (r & 1) == 1; k = first + ((r + 1) >> 1); }
                    if (k < last) {
                        if (isz) tz[at(k)] = x; else tq[at(k)] = x;
                    }
                    if (r == 0) sX[0] = x;
                    if (r == sn) sX[1] = x;
                    if (r == ru) sX[2] = x;
                }
            }
        } else if (lane == 0) {
            /* no system (one-cell run at an open start edge): the cell's
             * preindex is still needed at the end */
            Pre a = isU ? pre_uz(d, last, line) : pre_vz(d, line, last);
            sE[3] = a.a2; sE[4] = a.c2; sE[5] = a.d2;
        }
        __syncthreads();
        if (lane == 0) {
            if (first < last) {
                if (br1) tq[at(first - 1)] = 0;
                else if (dkQ1 && first == 2) tq[at(first - 1)] = ub1;
                else if (sn >= 0) tq[at(first - 1)] = (sE[2] - TZ_l - sE[1] * sX[0]) / sE[0];
            }
            if (br2) {
                tq[at(last)] = 0;
                if (sn >= 0) tz[at(last)] = sX[1];
            } else if (dkQ2 && last == dkfr) {
                tq[at(last)] = ub2;
                if (sn >= 0) tz[at(last)] = sX[1];
            } else {
                double prev = (first < last && sn >= 0) ? sX[2] : tq[at(last - 1)];
                tq[at(last)] = (sE[5] - sE[3] * prev - TZ_r) / sE[4];
            }
        }
        __syncthreads();
    }
}

__device__ __forceinline__ Pre shfl_pre(const Pre& a, int src)
{
    Pre b;
    b.f1 = __shfl(a.f1, src, kWave); b.f2 = __shfl(a.f2, src, kWave); b.f3 = __shfl(a.f3, src, kWave);
    b.f5 = __shfl(a.f5, src, kWave); b.a2 = __shfl(a.a2, src, kWave); b.c2 = __shfl(a.c2, src, kWave);
    b.d2 = __shfl(a.d2, src, kWave);
    return b;
}

/* Fused UZ / VZ solve of one wet run per wave with the speculative parallel
 * Thomas of hd_thomas.h: the same rows as k_adi_run (_calculate_preindex,
 * _calculate_abcd, _calculate_matrix_coeff, tridiag, _extract_solution,
 * UVZSolver_multithread.cu:203-723) and the same bits.  Chunks of 255 cells
 * (510 rows) are assembled by the 64 lanes, four cells each (cell t = lane +
 * 64 h; the preindex of cell t + 1 comes from the neighbouring lane by a
 * shuffle), into LDS; the forward sweep (factors replacing C / D) and the back
 * substitution (x replacing A) run over the chunk on all lanes; factors of
 * all but the last chunk go to the run's slice of the Ap/Bp line buffers and
 * come back for back substitution.  The run's wall-flag depths, boundary
 * depths, edge values and the first cells' preindex inputs are loaded
 * together, so a short run costs one memory round trip after its table
 * entry. */
/* repaired chunks of a sweep above which the next sweep of the same
 * direction takes 8 rows per lane */
constexpr int kAdaptRuns = 32;

template <bool isU>
__global__ void __launch_bounds__(64, 2) k_adi_spec(const HdDev* __restrict__ dv, int shift, int bnd)
{
    const HdDevG& d = dev_view(dv);
    const int sidx = 2 + shift, eidx = isU ? d.M - shift : d.N;
    constexpr int CH = 255; /* cells per LDS chunk (the rows of cell t need cells t and t + 1) */
    constexpr int RW = 2 * CH;
    __shared__ double sA[RW], sB[RW], sC[RW], sD[RW];
    __shared__ double sE[6];
    __shared__ double sX[3]; /* X(0), X(sn), X(last interior u row) */
    const int lane = threadIdx.x;
    const int N = d.N, M = d.M, W = d.W;
    const int tw = 2 * (N > M ? N : M) + 4;
    const int cap = (isU ? M + 3 : N + 3) * d.SL;
    const int cnt = d.run_cnt[isU ? 0 : 1] < cap ? d.run_cnt[isU ? 0 : 1] : cap;
    const int4* runs = isU ? d.runs_u : d.runs_v;
    const bool dkQ1 = isU ? d.bienQ[2] != 0 : d.bienQ[1] != 0;
    const bool dkQ2 = isU ? d.bienQ[3] != 0 : d.bienQ[0] != 0;
    const int dkfr = isU ? N : M;
    const double f4 = isU ? 2 * d.P.g * d.P.dTchia2dX : 2 * d.P.g * d.P.dTchia2dY;
    const double ND = d.P.NANGDAY;
    double* tz = d.t_z;
    double* tq = isU ? d.t_u : d.t_v;
    const double* hh = d.h;
    const int nbr = (lane + 1) & (kWave - 1);
    /* rows per lane of the speculative forward sweep: 8 while the previous
     * sweep of this direction repaired many runs (a flow starting up), else
     * as few as the run needs; this sweep's repaired chunks are counted for
     * the next one (moved into place by the explicit pass that follows) */
    const int minL = d.run_cnt[isU ? 2 : 3] > kAdaptRuns ? kSpecMaxL : 1;
    int* fails = d.flags + (isU ? 4 : 5);
    /* the extra last workgroup: the across edges' cells that no run holds */
    if (bnd && (int)blockIdx.x == (int)gridDim.x - 1) {
        across_boundary<isU>(d, edge_runs(d), lane);
        return;
    }
    const RunSlice sl = run_slice(cnt, (int)gridDim.x - bnd);
    for (int ri = sl.start; ri < sl.end; ri += sl.step) {
        const int4 rn = runs[ri];
        const int line = rn.x, first = rn.y, last = rn.z;
        if (line < sidx || line > eidx) continue;
        auto at = [&](int k) -> size_t { return isU ? (size_t)k * W + line : (size_t)line * W + k; };
        /* one round of loads: wall-flag depths (locate_segment_u / _v,
         * constant.cuh:135-178), boundary depths, edge values, first cells */
        double hb0, hb1, he0, he1;
        if (isU) {
            hb0 = hh[1 * W + line]; hb1 = hh[1 * W + line - 1];
            he0 = hh[N * W + line]; he1 = hh[N * W + line - 1];
        } else {
            hb0 = hh[line * W + first - 1]; hb1 = hh[(line - 1) * W + first - 1];
            he0 = hh[line * W + last]; he1 = hh[(line - 1) * W + last];
        }
        /* this run's boundary cells, before its end values are read (one
         * wave: its own stores are seen by its later loads) */
        if (bnd) run_boundary<isU>(d, line, first, last, lane);
        const double TZ_l = tz[at(first)];
        const double TZ_r = tz[at(last)];
        const double ub1 = isU ? d.ubt[line] : d.vbd[line];
        const double ub2 = isU ? d.ubp[line] : d.vbt[line];
        PreRaw raw0;
        {
            const int k = first + lane <= last ? first + lane : last;
            raw0 = isU ? pre_uz_ld(d, k, line) : pre_vz_ld(d, line, k);
        }
        const bool br1 = (first > 2) || ((first == 2) && ((hb0 + hb1) * 0.5 == ND));
        const bool br2 = isU ? (last < N) || ((last == N) && ((he0 + he1) * 0.5 == ND))
                             : (last < M) || ((last == M) && ((he0 + he1) * 0.5 == ND));
        int skind, ekind, sn;
        run_kinds(br1, br2, dkQ1, dkQ2, first, last, dkfr, skind, ekind, sn);
        const int ru = (skind != S_OPEN) ? 2 * (last - 1 - first) + 1 : 2 * (last - 1 - first);
        double* gAp = d.Ap + (size_t)line * tw + 2 * first;
        double* gBp = d.Bp + (size_t)line * tw + 2 * first;
        if (sn >= 0) {
            const int nch = sn / RW + 1;
            double ap = 0, bp = 0, xs = 0;
            for (int c = 0; c < nch; ++c) {
                const int c0 = first + CH * c;
                const int rb = RW * c;
                const int re = (rb + RW - 1 < sn) ? rb + RW - 1 : sn;
                /* cells t = lane + 64 h; the loads of h + 1 are issued before
                 * the rows of h are computed */
                auto eval_cell = [&](int h, bool have_raw) -> Pre {
                    const int k = c0 + lane + kWave * h;
                    PreRaw r;
                    if (have_raw) r = raw0;
                    else {
                        const int kc = k <= last ? k : last;
                        r = isU ? pre_uz_ld(d, kc, line) : pre_vz_ld(d, line, kc);
                    }
                    const Pre a = isU ? pre_uz_ev(d.P, r) : pre_vz_ev(d.P, r);
                    if (k == first) { sE[0] = a.a2; sE[1] = a.c2; sE[2] = a.d2; }
                    if (k == last) { sE[3] = a.a2; sE[4] = a.c2; sE[5] = a.d2; }
                    return a;
                };
                Pre cur = eval_cell(0, c == 0);
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    if (c0 + kWave * h > last) break; /* uniform */
                    const bool more = h < 3 && c0 + kWave * (h + 1) <= last;
                    Pre nxt = cur;
                    if (more) nxt = eval_cell(h + 1, false);
                    const Pre nA = shfl_pre(cur, nbr); /* cell t + 1 of lanes 0 .. 62 */
                    const Pre nB = shfl_pre(nxt, 0);   /* cell t + 1 of lane 63 */
                    const int t = lane + kWave * h;
                    const int k = c0 + t;
                    if (t < CH && k <= last) {
                        Pre b = cur;
                        if (k + 1 <= last) b = (lane < kWave - 1) ? nA : nB;
                        for (int e = 0; e < 2; ++e) {
                            int rr = 2 * (k - first) + e;
                            if (rr > sn) break;
                            Row4 w = cell_row(cur, b, f4, k, e, first, last, br1, br2, skind, ekind, sn, ub1, ub2, TZ_l,
                                              TZ_r);
                            int ix = rr - rb;
                            sA[ix] = w.A; sB[ix] = w.B; sC[ix] = w.C; sD[ix] = w.D;
                        }
                    }
                    cur = nxt;
                }
                __syncthreads();
                int g0 = 0;
                if (c == 0) {
                    if (sn == 0) {
                        xs = sD[0] / sB[0];
                    } else {
                        ap = -sC[0] / sB[0];
                        bp = sD[0] / sB[0];
                    }
                    g0 = 1;
                }
                const int g1 = (re < sn - 1 ? re : sn - 1) - rb;
                spec_forward<true>(sA, sB, sC, sD, g0, g1, ap, bp, minL, fails);
                if (c == 0 && sn > 0) {
                    __syncthreads();
                    if (lane == 0) { sC[0] = -sC[0] / sB[0]; sD[0] = sD[0] / sB[0]; }
                }
                if (re == sn && sn > 0) {
                    double A = sA[sn - rb], B = sB[sn - rb], D = sD[sn - rb];
                    xs = (D - (A * bp)) / (B + (A * ap));
                }
                __syncthreads();
                if (c + 1 < nch) { /* factors of this chunk (rows rb .. re, all < sn) to the line buffers */
                    for (int i = lane; i <= re - rb; i += kWave) { gAp[rb + i] = sC[i]; gBp[rb + i] = sD[i]; }
                    __syncthreads();
                }
            }
            /* back substitution, last chunk first; x[r] replaces A[r] */
            double xc = xs;
            for (int c = nch - 1; c >= 0; --c) {
                const int rb = RW * c;
                const int re = (rb + RW - 1 < sn) ? rb + RW - 1 : sn;
                if (c + 1 < nch) {
                    __syncthreads();
                    for (int i = lane; i <= re - rb; i += kWave) { sC[i] = gAp[rb + i]; sD[i] = gBp[rb + i]; }
                    __syncthreads();
                } else if (lane == 0) {
                    sA[sn - rb] = xs;
                }
                const int hi = (re < sn - 1 ? re : sn - 1) - rb;
                xc = spec_backward(sC, sD, sA, hi, xc);
                __syncthreads();
                /* scatter the unknowns of positions first .. last-1 */
                for (int i = lane; i <= re - rb; i += kWave) {
                    int r = rb + i;
                    double x = sA[i];
                    int k;
                    bool isz;
                    if (skind != S_OPEN) { k = first + (r >> 1); isz = (r & 1) == 0; }
                    else { isz = (r & 1) == 1; k = first + ((r + 1) >> 1); }
                    if (k < last) {
                        if (isz) tz[at(k)] = x; else tq[at(k)] = x;
                    }
                    if (r == 0) sX[0] = x;
                    if (r == sn) sX[1] = x;
                    if (r == ru) sX[2] = x;
                }
            }
        } else {
            /* no system (one-cell run at an open start edge): the cell's
             * preindex is still needed at the end (first == last: lane 0
             * loaded it) */
            if (lane == 0) {
                Pre a = isU ? pre_uz_ev(d.P, raw0) : pre_vz_ev(d.P, raw0);
                sE[3] = a.a2; sE[4] = a.c2; sE[5] = a.d2;
            }
        }
        __syncthreads();
        if (lane == 0) {
            if (first < last) {
                if (br1) tq[at(first - 1)] = 0;
                else if (dkQ1 && first == 2) tq[at(first - 1)] = ub1;
                else if (sn >= 0) tq[at(first - 1)] = (sE[2] - TZ_l - sE[1] * sX[0]) / sE[0];
            }
            if (br2) {
                tq[at(last)] = 0;
                if (sn >= 0) tz[at(last)] = sX[1];
            } else if (dkQ2 && last == dkfr) {
                tq[at(last)] = ub2;
                if (sn >= 0) tz[at(last)] = sX[1];
            } else {
                double prev = (first < last && sn >= 0) ? sX[2] : tq[at(last - 1)];
                tq[at(last)] = (sE[5] - sE[3] * prev - TZ_r) / sE[4];
            }
        }
        __syncthreads();
    }
}

/* Cell (i, j) of a thread of a per-row kernel: with `tiled` the blocks walk
 * the context's list of active 64-column row tiles (tiles that hold, or
 * border, a cell with khouot != 2: elsewhere the per-row kernels are at a
 * fixed point once a full step has run), else the full padded grid
 * (blockIdx.y = row).  Returns false for threads outside the grid. */
__device__ __forceinline__ bool row_cell(const HdDevG& d, int tiled, int& i, int& j)
{
    if (tiled) {
        int t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
        if (t >= d.ntiles[0]) return false;
        int tile = d.tiles[t];
        int ncb = (d.M + 3 + 63) >> 6;
        i = tile / ncb;
        j = (tile % ncb) * 64 + (threadIdx.x & 63);
    } else {
        j = blockIdx.x * blockDim.x + threadIdx.x;
        i = blockIdx.y;
    }
    return j < d.M + 3 && i < d.N + 3;
}

/* ================================================================ smoothing */

/* Normalize + update_buffer, support_funcs.cu:682-729, writing the partner
 * buffer (the caller swaps them).  With `post` set it also does
 * update_h_moi (support_funcs.cu:60-73), the drying test of Reset_states
 * (support_funcs.cu:114-127 / 173-185) for every cell, and records the
 * positions where the serial wetting pass can act (a neighbour along the
 * reset direction has khouot == 1). */
__global__ void __launch_bounds__(256) k_normalize(const HdDev* __restrict__ dv, int isU, int post, int horizontal,
                                                    int tiled)
{
    const HdDevG& d = dev_view(dv);
    int* cand_cnt = d.cand_cnt;
    int* cand = d.cand;
    unsigned char* dryflag = d.dryflag;
    int i, j;
    if (!row_cell(d, tiled, i, j)) return;
    int N = d.N, M = d.M, W = d.W;
    int p = i * W + j;
    const double* val = isU ? d.t_u : d.t_v;
    double* out = isU ? d.tmp_u : d.tmp_v;
    int close = isU ? W : 1, far = isU ? 1 : W;
    /* every value the cell may read, in one round of loads (neighbour
     * indices fall back to the cell itself outside the ranges that use them) */
    const bool in = i >= 2 && i <= N && j >= 2 && j <= M;
    const double r0 = val[p];
    const double vcm = val[in ? p - close : p], vcp = val[in ? p + close : p];
    const double vfm = val[in ? p - far : p], vfp = val[in ? p + far : p];
    const int kfm = d.khouot[in ? p - far : p], kfp = d.khouot[in ? p + far : p];
    int kh = 0, kcm = 0, kcp = 0;
    double htz = 0, tz = 0, zz = 0;
    bool cin = false;
    if (post) {
        kh = d.khouot[p];
        htz = d.htaiz[p];
        tz = d.t_z[p];
        zz = d.z[p];
        cin = horizontal ? (i >= 1 && i <= N && j >= 2 && j <= M) : (j >= 1 && j <= M && i >= 2 && i <= N);
        const int step = horizontal ? 1 : W;
        kcm = d.khouot[cin ? p - step : p];
        kcp = d.khouot[cin ? p + step : p];
    }
    double r = r0;
    if (in && r != 0) {
        int count = 2;
        double nb = vcm + vcp;
        if (kfm == 0) { nb += vfm; count++; }
        if (kfp == 0) { nb += vfp; count++; }
        r = r * d.P.heso + (1 - d.P.heso) * nb / count;
    }
    out[p] = r;
    if (!post) return;
    unsigned char dry = 0;
    if (i >= 1 && i <= N && j >= 1 && j <= M && kh == 0) {
        double hm = htz + tz;
        d.H_moi[p] = hm;
        bool inrange = horizontal ? (j >= 2) : (i >= 2);
        dry = inrange && (hm <= d.P.H_TINH);
    }
    dryflag[p] = dry;
    {   /* the wet bits of this 64-cell tile row (the wave) */
        const unsigned long long bits = __ballot(kh == 0 && !dry);
        if ((threadIdx.x & 63) == 0) d.wetb[(size_t)i * ((M + 3 + 63) >> 6) + (j >> 6)] = bits;
    }
    /* a position can open a neighbour only where t_z > z; t_z changes
     * during the wetting pass only at cells it opens, and those are reached
     * by the pass's own cascade */
    if (cin && tz > zz && (kcm == 1 || kcp == 1)) {
        if (horizontal) {
            int slot = atomicAdd(&cand_cnt[i], 1);
            if (slot < kCand) cand[i * kCand + slot] = j;
        } else {
            int slot = atomicAdd(&cand_cnt[j], 1);
            if (slot < kCand) cand[j * kCand + slot] = i;
        }
    }
}

/* ================================================================ explicit */

/* Every value vSolver / uSolver may read at one cell, loaded
 * unconditionally (all in bounds for the cells the explicit kernels visit)
 * so that the loads of several cells are in flight together; the
 * evaluations below take the reference's branches on them. */
struct ExRaw {
    double c, a1, b1;              /* the velocity at the cell and at +-1 along its own direction */
    double n1, n2, m1, m2;         /* the velocity at +1, +2, -1, -2 across */
    double x0, x1, x2, x3;         /* the other velocity at the four surrounding faces */
    double t0, t1, t2, t3;         /* the same from its new (smoothed) field */
    double ht, k1, hm_m1, hm_p1, hm_m2, zf, zc, vis, ts;
};

/* loads of vSolver (UVZSolver_multithread.cu:727-768) at (row, col) */
__device__ __forceinline__ ExRaw v_solver_ld(const HdDevG& d, int row, int col)
{
    const int W = d.W, p = row * W + col;
    ExRaw r;
    r.c = d.v[p]; r.a1 = d.v[p + 1]; r.b1 = d.v[p - 1];
    r.n1 = d.v[p + W]; r.n2 = d.v[p + 2 * W]; r.m1 = d.v[p - W]; r.m2 = d.v[p - 2 * W];
    r.x0 = d.u[p - W]; r.x1 = d.u[p]; r.x2 = d.u[p - W + 1]; r.x3 = d.u[p + 1];
    r.t0 = d.t_u[p - W]; r.t1 = d.t_u[p]; r.t2 = d.t_u[p - W + 1]; r.t3 = d.t_u[p + 1];
    r.ht = d.Htdv[p]; r.k1 = d.Ky1[p];
    r.hm_m1 = d.H_moi[p - W]; r.hm_p1 = d.H_moi[p + W]; r.hm_m2 = d.H_moi[p - 2 * W];
    r.zf = d.z[p + 1]; r.zc = d.z[p]; r.vis = d.VISCOIDX[p]; r.ts = d.Tsyw[p];
    return r;
}

/* vSolver, UVZSolver_multithread.cu:727-768, on the loaded values; reads the
 * smoothed t_u. */
__device__ __forceinline__ double v_solver_ev(const hd_params& P, const ExRaw& r)
{
    double p, q, tmp;
    q = 0; p = 0; tmp = 0;
    const double vc = r.c;
    const double utb = (r.x0 + r.x1 + r.x2 + r.x3) * 0.25;
    const double t_utb = (r.t0 + r.t1 + r.t2 + r.t3) * 0.25;
    const double htv = r.ht;
    p = (r.a1 - r.b1) / P.dY2;
    p = P.HaiChiadT + p + r.k1 * sqrt(utb * utb + vc * vc) / htv;
    if (r.hm_m1 <= P.H_TINH) {
        if (utb < 0) {
            q = t_utb * (-3 * vc + 4 * r.n1 - r.n2) / P.dX2;
            tmp = (vc - 2 * r.n1 + r.n2) / P.dXbp;
        }
    } else {
        if (r.hm_p1 <= P.H_TINH) {
            if ((r.hm_m2 > P.H_TINH) && (utb > 0)) {
                q = t_utb * (3 * vc - 4 * r.m1 + r.m2) / P.dX2;
                tmp = (vc - 2 * r.m1 + r.m2) / P.dXbp;
            }
        } else {
            q = t_utb * (r.n1 - r.m1) / P.dX2;
            tmp = (r.n1 - 2 * vc + r.m1) / P.dXbp;
        }
    }
    q = P.HaiChiadT * vc - q - P.CORIOLIS_FORCE * t_utb;
    q = (q - P.g * (r.zf - r.zc) / P.dY + r.vis * (tmp + (r.a1 - 2 * vc + r.b1) / P.dYbp)) + (P.Windy - r.ts) / htv;
    return q / p;
}

/* loads of uSolver (UVZSolver_multithread.cu:773-815) at (row, col) */
__device__ __forceinline__ ExRaw u_solver_ld(const HdDevG& d, int row, int col)
{
    const int W = d.W, p = row * W + col;
    ExRaw r;
    r.c = d.u[p]; r.a1 = d.u[p + W]; r.b1 = d.u[p - W];
    r.n1 = d.u[p + 1]; r.n2 = d.u[p + 2]; r.m1 = d.u[p - 1]; r.m2 = d.u[p - 2];
    r.x0 = d.v[p - 1]; r.x1 = d.v[p]; r.x2 = d.v[p + W - 1]; r.x3 = d.v[p + W];
    r.t0 = d.t_v[p - 1]; r.t1 = d.t_v[p]; r.t2 = d.t_v[p + W - 1]; r.t3 = d.t_v[p + W];
    r.ht = d.Htdu[p]; r.k1 = d.Kx1[p];
    r.hm_m1 = d.H_moi[p - 1]; r.hm_p1 = d.H_moi[p + 1]; r.hm_m2 = d.H_moi[p - 2];
    r.zf = d.z[p + W]; r.zc = d.z[p]; r.vis = d.VISCOIDX[p]; r.ts = d.Tsxw[p];
    return r;
}

/* uSolver, UVZSolver_multithread.cu:773-815, on the loaded values; reads the
 * smoothed t_v. */
__device__ __forceinline__ double u_solver_ev(const hd_params& P, const ExRaw& r)
{
    double p, q, tmp;
    p = 0; q = 0; tmp = 0;
    const double uc = r.c;
    const double vtb = (r.x0 + r.x1 + r.x2 + r.x3) * 0.25;
    const double t_vtb = (r.t0 + r.t1 + r.t2 + r.t3) * 0.25;
    const double htu = r.ht;
    p = (r.a1 - r.b1) / P.dX2;
    p = (P.HaiChiadT + p + r.k1 * sqrt(vtb * vtb + uc * uc) / htu);
    if (r.hm_m1 <= P.H_TINH) {
        if (vtb < 0) {
            q = t_vtb * (-3 * uc + 4 * r.n1 - r.n2) / P.dY2;
            tmp = (uc - 2 * r.n1 + r.n2) / P.dYbp;
        }
    } else {
        if (r.hm_p1 <= P.H_TINH) {
            if ((r.hm_m2 > P.H_TINH) && (vtb > 0)) {
                q = t_vtb * (3 * uc - 4 * r.m1 + r.m2) / P.dY2;
                tmp = (uc - 2 * r.m1 + r.m2) / P.dYbp;
            }
        } else {
            q = t_vtb * (r.n1 - r.m1) / P.dY2;
            tmp = (r.n1 - 2 * uc + r.m1) / P.dYbp;
        }
    }
    q = P.HaiChiadT * uc - q + P.CORIOLIS_FORCE * t_vtb;
    q = (q - P.g * (r.zf - r.zc) / P.dX + r.vis * (tmp + (r.a1 - 2 * uc + r.b1) / P.dXbp)) + (P.Windx - r.ts) / htu;
    return q / p;
}

__device__ double v_solver(const HdDevG& d, int row, int col) { return v_solver_ev(d.P, v_solver_ld(d, row, col)); }
__device__ double u_solver(const HdDevG& d, int row, int col) { return u_solver_ev(d.P, u_solver_ld(d, row, col)); }

/* solveV (UVZSolver_multithread.cu:1055-1071) or solveU (:947-968): one
 * thread per cell, the cell's run from an LDS copy of the line tables; cells
 * first <= k < last of runs with first < last. */
template <bool isU>
__global__ void __launch_bounds__(256) k_explicit(const HdDev* __restrict__ dv)
{
    const HdDevG& d = dev_view(dv);
    const int sidx = 2, eidx = isU ? d.M : d.N;
    __shared__ int s_moc[64];
    __shared__ int s_dau[64][24];
    __shared__ int s_cuoi[64][24];
    int N = d.N, M = d.M, SL = d.SL;
    int tx = threadIdx.x, ty = threadIdx.y;
    int j0 = blockIdx.x * 64, i0 = blockIdx.y * 16;
    int tid = ty * 64 + tx;
    /* solveU: column lines (dauj); solveV: row lines (daui) */
    if (isU) {
        for (int t = tid; t < 64; t += 256) {
            int L = j0 + t;
            s_moc[t] = (L >= sidx && L <= eidx && L <= M + 2) ? min(d.mocj[L], SL) : 0;
        }
        __syncthreads();
        for (int t = tid; t < 64 * SL; t += 256) {
            int c = t / SL, s = t % SL;
            if (s < s_moc[c]) { s_dau[c][s] = d.dauj[(j0 + c) * SL + s]; s_cuoi[c][s] = d.cuoij[(j0 + c) * SL + s]; }
        }
    } else {
        for (int t = tid; t < 16; t += 256) {
            int L = i0 + t;
            s_moc[t] = (L >= sidx && L <= eidx && L <= N + 2) ? min(d.moci[L], SL) : 0;
        }
        __syncthreads();
        for (int t = tid; t < 16 * SL; t += 256) {
            int c = t / SL, s = t % SL;
            if (s < s_moc[c]) { s_dau[c][s] = d.daui[(i0 + c) * SL + s]; s_cuoi[c][s] = d.cuoii[(i0 + c) * SL + s]; }
        }
    }
    __syncthreads();
    int j = j0 + tx;
    if (j > M) return;
    for (int r = ty; r < 16; r += 4) {
        int i = i0 + r;
        if (isU) { if (i < 1 || i > N) continue; }
        else { if (i > N) break; }
        int li = isU ? tx : r;
        int k = isU ? i : j;
        int first, last;
        if (find_run(s_dau[li], s_cuoi[li], s_moc[li], k, first, last) < 0) continue;
        if (isU) {
            if (first > last || k < first - 1 || k >= last) continue;
            d.t_u[i * d.W + j] = u_solver(d, i, j);
        } else {
            if (first >= last || k >= last || k < first) continue;
            d.t_v[i * d.W + j] = v_solver(d, i, j);
        }
    }
}

/* update_margin_elem_V (UVZSolver_multithread.cu:1008-1052) or
 * update_margin_elem_U (:970-1006): one lane per line, runs in order. */
template <bool isU>
__global__ void k_margins(const HdDev* __restrict__ dv)
{
    const HdDevG& d = dev_view(dv);
    const int sidx = 2, eidx = isU ? d.M : d.N;
    int line = sidx + blockIdx.x * blockDim.x + threadIdx.x;
    if (line > eidx) return;
    int N = d.N, M = d.M, W = d.W, SL = d.SL;
    const double* h = d.h;
    const double eps = 1e-16; /* UVZSolver_multithread.cu:12 */
    double ND = d.P.NANGDAY;
    int nseg = isU ? d.mocj[line] : d.moci[line];
    if (nseg > SL) nseg = SL;
    for (int k = 0; k < nseg; ++k) {
        if (isU) {
            int j = line;
            int first = d.dauj[j * SL + k], last = d.cuoij[j * SL + k];
            bool br1 = (first > 2) || ((first == 2) && ((fabs(h[1 * W + j] + h[1 * W + j - 1]) * 0.5 - ND) < eps));
            bool br2 = (last < N) || ((last == N) && (fabs((h[N * W + j] + h[N * W + j - 1]) * 0.5 - ND) < eps));
            double* t = d.t_u;
            if (br1) t[(first - 1) * W + j] = 0;
            else t[(first - 1) * W + j] = 2 * t[first * W + j] - t[(first + 1) * W + j];
            if (br2) t[last * W + j] = 0;
            else t[last * W + j] = 2 * t[(last - 1) * W + j] - t[(last - 2) * W + j];
        } else {
            int i = line;
            int first = d.daui[i * SL + k], last = d.cuoii[i * SL + k];
            bool br1 = (first > 2) || ((first == 2) && (fabs((h[i * W + first - 1] + h[(i - 1) * W + first - 1]) * 0.5 - ND) < eps));
            bool br2 = (last < M) || ((last == M) && (fabs((h[i * W + last] + h[(i - 1) * W + last]) * 0.5 - ND) < eps));
            double* t = d.t_v;
            if (br1) t[i * W + first - 1] = 0;
            else t[i * W + first - 1] = 2 * t[i * W + first] - t[i * W + first + 1];
            if (br2) t[i * W + last] = 0;
            else t[i * W + last] = 2 * t[i * W + last - 1] - t[i * W + last - 2];
        }
    }
}

/* solveU / solveV followed by the margins of the same line, one wave per
 * line: the wave computes the explicit velocity of every cell first <= k <
 * last of each wet run of the line (lanes over cells), then, after a
 * workgroup barrier, lane 0 applies update_margin_elem_U / _V to the runs in
 * order.  Explicit velocities read only the other direction's half-step
 * field and the old state, and a line's margins read only that line, so this
 * is the two-kernel order (all explicit cells, then margins) line by line. */
template <bool isU>
__global__ void __launch_bounds__(64) k_explicit_margins(const HdDev* __restrict__ dv, int cells)
{
    const HdDevG& d = dev_view(dv);
    const int eidx = isU ? d.M : d.N;
    const int line = 2 + blockIdx.x;
    if (line > eidx) return;
    const int N = d.N, M = d.M, W = d.W, SL = d.SL;
    const int lane = threadIdx.x;
    int nseg = isU ? d.mocj[line] : d.moci[line];
    if (nseg > SL) nseg = SL;
    const int* dau = (isU ? d.dauj : d.daui) + line * SL;
    const int* cuoi = (isU ? d.cuoij : d.cuoii) + line * SL;
    /* the cells first .. last-1 of all runs of the line, numbered across the
     * runs (run s starts at s_ex[s]); four cells per lane per round, all
     * loads of a round in flight before any store */
    __shared__ int s_ex[25], s_f[24];
    int len = 0, f0 = 0;
    if (lane < nseg) {
        f0 = dau[lane];
        const int l0 = cuoi[lane];
        len = l0 > f0 ? l0 - f0 : 0;
    }
    int incl = len;
    for (int dd = 1; dd < 32; dd <<= 1) {
        const int v = __shfl_up(incl, dd, kWave);
        if (lane >= dd) incl += v;
    }
    if (lane < nseg) { s_ex[lane] = incl - len; s_f[lane] = f0; }
    const int total = __shfl(incl, nseg > 0 ? nseg - 1 : 0, kWave) * (nseg > 0);
    __syncthreads();
    constexpr int B = 4;
    for (int base = 0; cells && base < total; base += B * kWave) {
        int kk[B];
        bool on[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int idx = base + lane + b * kWave;
            int sg = 0;
            while (sg + 1 < nseg && s_ex[sg + 1] <= idx) ++sg;
            const int k = s_f[sg] + (idx - s_ex[sg]);
            on[b] = idx < total && (isU ? (k >= 1 && k <= N) : (k <= M && line <= N));
            kk[b] = on[b] ? k : s_f[0]; /* an inactive slot loads a cell of the line's first run */
        }
        ExRaw r[B];
#pragma unroll
        for (int b = 0; b < B; ++b) r[b] = isU ? u_solver_ld(d, kk[b], line) : v_solver_ld(d, line, kk[b]);
#pragma unroll
        for (int b = 0; b < B; ++b) {
            if (!on[b]) continue;
            if (isU) d.t_u[kk[b] * W + line] = u_solver_ev(d.P, r[b]);
            else d.t_v[line * W + kk[b]] = v_solver_ev(d.P, r[b]);
        }
    }
    __syncthreads();
    if (lane != 0) return;
    const double* h = d.h;
    const double eps = 1e-16; /* UVZSolver_multithread.cu:12 */
    const double ND = d.P.NANGDAY;
    for (int sg = 0; sg < nseg; ++sg) {
        const int first = dau[sg], last = cuoi[sg];
        if (isU) {
            const int j = line;
            bool br1 = (first > 2) || ((first == 2) && ((fabs(h[1 * W + j] + h[1 * W + j - 1]) * 0.5 - ND) < eps));
            bool br2 = (last < N) || ((last == N) && (fabs((h[N * W + j] + h[N * W + j - 1]) * 0.5 - ND) < eps));
            double* t = d.t_u;
            if (br1) t[(first - 1) * W + j] = 0;
            else t[(first - 1) * W + j] = 2 * t[first * W + j] - t[(first + 1) * W + j];
            if (br2) t[last * W + j] = 0;
            else t[last * W + j] = 2 * t[(last - 1) * W + j] - t[(last - 2) * W + j];
        } else {
            const int i = line;
            bool br1 = (first > 2) || ((first == 2) && (fabs((h[i * W + first - 1] + h[(i - 1) * W + first - 1]) * 0.5 - ND) < eps));
            bool br2 = (last < M) || ((last == M) && (fabs((h[i * W + last] + h[(i - 1) * W + last]) * 0.5 - ND) < eps));
            double* t = d.t_v;
            if (br1) t[i * W + first - 1] = 0;
            else t[i * W + first - 1] = 2 * t[i * W + first] - t[i * W + first + 1];
            if (br2) t[i * W + last] = 0;
            else t[i * W + last] = 2 * t[i * W + last - 1] - t[i * W + last - 2];
        }
    }
}

/* solveU (UVZSolver_multithread.cu:947-968) over the active row tiles: the
 * cells first .. last-1 of every wet run of every column.  They are
 * recognised from the mask the run tables were built from (khouot changes
 * only later in the half-step): along a column of positions 2 .. N, cell k
 * belongs to a run's first .. last-1 iff k and k+1 are wet (khouot == 0) and
 * k's wet interval starts at or before N-2 -- for k <= N-2 always, for
 * k = N-1 iff N-2 is wet -- which is how Find_Calculation_limits_Vertical
 * (support_funcs.cu:252-279) forms the runs.  A wave's loads are then 64
 * adjacent columns of one row instead of 64 rows of one column;
 * k_explicit_margins (cells off) writes the margins after it. */
__global__ void __launch_bounds__(256) k_explicit_rows_u(const HdDev* __restrict__ dv, int tiled)
{
    const HdDevG& d = dev_view(dv);
    int i, j;
    if (!row_cell(d, tiled, i, j)) return;
    const int N = d.N, M = d.M, W = d.W;
    if (j < 2 || j > M || i < 2 || i > N - 1) return;
    const int* kh = d.khouot;
    const int p = i * W + j;
    const bool in = kh[p] == 0 && kh[p + W] == 0 && (i <= N - 2 || kh[(N - 2) * W + j] == 0);
    if (in) d.t_u[p] = u_solver(d, i, j);
}

/* Normalize of t_v (support_funcs.cu:682-703, the V form: neighbours along
 * the row always, across rows when wet) at cell (qi, qj), from the unsmoothed
 * buffer, with every load issued unconditionally.  The same operations as
 * k_normalize(isU = 0), so the value equals what that pass writes. */
__device__ __forceinline__ double smooth_v_at(const HdDevG& d, int qi, int qj)
{
    const int N = d.N, M = d.M, W = d.W;
    const int q = qi * W + qj;
    const bool in = qi >= 2 && qi <= N && qj >= 2 && qj <= M;
    const double* val = d.t_v;
    const double r0 = val[q];
    const double vcm = val[in ? q - 1 : q], vcp = val[in ? q + 1 : q];
    const double vfm = val[in ? q - W : q], vfp = val[in ? q + W : q];
    const int kfm = d.khouot[in ? q - W : q], kfp = d.khouot[in ? q + W : q];
    double r = r0;
    if (in && r != 0) {
        int count = 2;
        double nb = vcm + vcp;
        if (kfm == 0) { nb += vfm; count++; }
        if (kfp == 0) { nb += vfp; count++; }
        r = r * d.P.heso + (1 - d.P.heso) * nb / count;
    }
    return r;
}

/* Normalize of t_u (the U form: neighbours along the column always, across
 * columns when wet) at cell (qi, qj), as smooth_v_at. */
__device__ __forceinline__ double smooth_u_at(const HdDevG& d, int qi, int qj)
{
    const int N = d.N, M = d.M, W = d.W;
    const int q = qi * W + qj;
    const bool in = qi >= 2 && qi <= N && qj >= 2 && qj <= M;
    const double* val = d.t_u;
    const double r0 = val[q];
    const double vcm = val[in ? q - W : q], vcp = val[in ? q + W : q];
    const double vfm = val[in ? q - 1 : q], vfp = val[in ? q + 1 : q];
    const int kfm = d.khouot[in ? q - 1 : q], kfp = d.khouot[in ? q + 1 : q];
    double r = r0;
    if (in && r != 0) {
        int count = 2;
        double nb = vcm + vcp;
        if (kfm == 0) { nb += vfm; count++; }
        if (kfp == 0) { nb += vfp; count++; }
        r = r * d.P.heso + (1 - d.P.heso) * nb / count;
    }
    return r;
}

/* solveV's t_v at (i, x) / solveU's t_u at (x, j) as the fused kernels
 * below compute it: vSolver / uSolver on the other direction's smoothed
 * field at the four faces around the cell. */
__device__ __forceinline__ double explicit_v_at(const HdDevG& d, int i, int x)
{
    ExRaw r = v_solver_ld(d, i, x);
    r.t0 = smooth_u_at(d, i - 1, x);
    r.t1 = smooth_u_at(d, i, x);
    r.t2 = smooth_u_at(d, i - 1, x + 1);
    r.t3 = smooth_u_at(d, i, x + 1);
    return v_solver_ev(d.P, r);
}

__device__ __forceinline__ double explicit_u_at(const HdDevG& d, int x, int j)
{
    ExRaw r = u_solver_ld(d, x, j);
    r.t0 = smooth_v_at(d, x, j - 1);
    r.t1 = smooth_v_at(d, x, j);
    r.t2 = smooth_v_at(d, x + 1, j - 1);
    r.t3 = smooth_v_at(d, x + 1, j);
    return u_solver_ev(d.P, r);
}

/* update_margin_elem_V (UVZSolver_multithread.cu:1008-1052; ROWS: the runs
 * of row `line`, positions = columns, L = M) / update_margin_elem_U
 * (:970-1006; the runs of column `line`, positions = rows, L = N) at
 * position x, folded into the explicit pass.  The runs are those of the
 * segment tables, recognised from the mask they were built from (khouot
 * changes only later in the half-step); wx, wn, wl2, wl1 are the wetness of
 * positions x, x+1, L-2 and L-1 of the line.  A run [first, last] has
 * margins first-1 and last, both 0 except at an open grid edge: first == 2
 * with the depth at first-1 not at the wall value gives 2 E(2) - E(3),
 * last == L likewise 2 E(L-1) - E(L-2), E being the line after solveU /
 * solveV.  E(L-1), E(L-2) are explicit cells of the run, E(2), E(3) too
 * unless the run is [2, 2] or [2, 3], where the value before the margins
 * enters; the threads of positions 1 and L evaluate them (edge_vals, `ev`).
 * In those short cases the run's margin last (and, after [2, 2], the margin
 * 3 of a run starting at 4) is written by position 1's thread after it read
 * them, in the reference's order, and not by their own threads.  Returns the
 * number of margin values this thread writes (0..3), at positions pos[k]
 * with values val[k]; only positions 1..3 and L load anything. */
struct EdgeVals {
    double ea, eb, o2, o3; /* E(2), E(3), old 2, old 3 at position 1; E(L-1), E(L-2) at L */
};

/* The values an open-edge margin may use, evaluated with the pass's other
 * loads by the threads of positions 1 and L, which have no explicit cell of
 * their own.  The old values of positions 2 and 3 are used only where no
 * other thread writes those positions. */
template <bool ROWS>
__device__ __forceinline__ EdgeVals edge_vals(const HdDevG& d, int line, int x)
{
    EdgeVals e{0.0, 0.0, 0.0, 0.0};
    const int W = d.W, L = ROWS ? d.M : d.N;
    if (line < 2 || line > (ROWS ? d.N : d.M)) return e;
    auto at = [&](int k) { return ROWS ? line * W + k : k * W + line; };
    auto ex_at = [&](int k) -> double { return ROWS ? explicit_v_at(d, line, k) : explicit_u_at(d, k, line); };
    if (x == 1) {
        const double* t = ROWS ? d.t_v : d.t_u;
        e.o2 = t[at(2)];
        e.o3 = t[at(3)];
        e.ea = ex_at(2);
        e.eb = ex_at(3);
    } else if (x == L) {
        e.ea = ex_at(L - 1);
        e.eb = ex_at(L - 2);
    }
    return e;
}

template <bool ROWS>
__device__ __forceinline__ int margin_at(const HdDevG& d, int line, int x, bool wx, bool wn, bool wl2, bool wl1,
                                         const EdgeVals& ev, int pos[3], double val[3])
{
    const int N = d.N, M = d.M, W = d.W;
    const int L = ROWS ? M : N;
    if (x < 1 || x > L) return 0;
    const bool first_m = wn && !wx && x + 1 <= L - 2;
    const bool last_m = wx && !wn && (x <= L - 2 || (x == L - 1 && wl2) || (x == L && wl1 && wl2));
    if (!first_m && !last_m) return 0;
    pos[0] = x;
    val[0] = 0.0;
    if (x > 3 && x < L) return 1; /* first > 4, or last < L: a wall margin */
    const int* kh = d.khouot;
    const double* h = d.h;
    auto at = [&](int k) { return ROWS ? line * W + k : k * W + line; };
    auto wet = [&](int k) -> bool { return k >= 2 && k <= L && kh[at(k)] == 0; };
    const double eps = 1e-16; /* UVZSolver_multithread.cu:12 */
    const double ND = d.P.NANGDAY;
    if (x == L) {
        const bool wall2 = ROWS ? fabs((h[line * W + M] + h[(line - 1) * W + M]) * 0.5 - ND) < eps
                                : fabs((h[N * W + line] + h[N * W + line - 1]) * 0.5 - ND) < eps;
        if (!wall2) val[0] = 2 * ev.ea - ev.eb;
        return 1;
    }
    /* x <= 3: the run from 2, if any -- does it end at 2 or 3, does its
     * first margin extrapolate? */
    const bool wall1 = ROWS ? fabs((h[line * W + 1] + h[(line - 1) * W + 1]) * 0.5 - ND) < eps
                            : (fabs(h[1 * W + line] + h[1 * W + line - 1]) * 0.5 - ND) < eps;
    const bool w2 = wet(2), w3 = wet(3), w4 = wet(4);
    const bool end2 = w2 && !w3, end3 = w2 && w3 && !w4;
    if (w2 && !wall1 && (end2 || end3)) {
        if (x != 1) return 0;
        const double e2 = end2 ? ev.o2 : ev.ea;
        val[0] = 2 * e2 - ev.o3;
        pos[1] = end2 ? 2 : 3;
        val[1] = 0.0;
        if (end2 && w4 && 4 <= L - 2) { pos[2] = 3; val[2] = 0.0; return 3; }
        return 2;
    }
    if (x == 1 && !wall1) val[0] = 2 * ev.ea - ev.eb;
    return 1;
}

/* The drying test of Reset_states (support_funcs.cu:114-127 / 173-185) and
 * the wetting candidates at cell (i, j), as the drying pass (k_normalize with
 * post) computes them, but without the H_moi write (update_h_moi,
 * support_funcs.cu:60-73): the explicit cells of the same launch read the
 * old H_moi, and the smoothing + wetting launch that follows writes it.
 * dry_load does the loads, dry_store the drying flag, the wave's words of
 * the wet bits (khouot == 0 and not drying) and of the pre-wetting mask bits
 * kh0b (khouot == 0), and the candidate record; every lane of the wave that
 * holds a grid cell calls dry_store. */
struct DryMark {
    int kh, kcm, kcp;
    double htz, tz, zz;
};

__device__ __forceinline__ DryMark dry_load(const HdDevG& d, int i, int j, int horizontal)
{
    const int N = d.N, M = d.M, W = d.W, p = i * W + j;
    const bool cin = horizontal ? (i >= 1 && i <= N && j >= 2 && j <= M) : (j >= 1 && j <= M && i >= 2 && i <= N);
    const int step = horizontal ? 1 : W;
    DryMark m;
    m.kh = d.khouot[p];
    m.htz = d.htaiz[p];
    m.tz = d.t_z[p];
    m.zz = d.z[p];
    m.kcm = d.khouot[cin ? p - step : p];
    m.kcp = d.khouot[cin ? p + step : p];
    return m;
}

__device__ __forceinline__ void dry_store(const HdDevG& d, int i, int j, int horizontal, const DryMark& m)
{
    const int N = d.N, M = d.M, W = d.W, p = i * W + j;
    unsigned char dry = 0;
    if (i >= 1 && i <= N && j >= 1 && j <= M && m.kh == 0) {
        const double hm = m.htz + m.tz;
        const bool inrange = horizontal ? (j >= 2) : (i >= 2);
        dry = inrange && (hm <= d.P.H_TINH);
    }
    d.dryflag[p] = dry;
    const unsigned long long wb = __ballot(m.kh == 0 && !dry), kb = __ballot(m.kh == 0);
    if ((threadIdx.x & 63) == 0) {
        const size_t w = (size_t)i * ((M + 3 + 63) >> 6) + (j >> 6);
        d.wetb[w] = wb;
        d.kh0b[w] = kb;
    }
    const bool cin = horizontal ? (i >= 1 && i <= N && j >= 2 && j <= M) : (j >= 1 && j <= M && i >= 2 && i <= N);
    if (cin && m.tz > m.zz && (m.kcm == 1 || m.kcp == 1)) {
        const int ln = horizontal ? i : j, pos = horizontal ? j : i;
        const int slot = atomicAdd(&d.cand_cnt[ln], 1);
        if (slot < kCand) d.cand[ln * kCand + slot] = pos;
    }
}

/* The UZ half-step's smoothing of t_u and solveV (UVZSolver_multithread.cu:
 * 1055-1071) in one pass over the active row tiles: the cells first ..
 * last-1 of every wet run of every row, recognised from the mask as in
 * k_explicit_rows_u (k and k+1 wet, and for k = M-1 also M-2), with the
 * smoothed t_u at the four faces around the cell recomputed from the
 * unsmoothed buffer, the margins of the rows' runs (margin_at), and the
 * drying flags, wet bits and wetting candidates of the row reset
 * (dry_load / dry_store). */
__global__ void __launch_bounds__(256) k_smooth_u_explicit_v(const HdDev* __restrict__ dv, int tiled, int adv)
{
    const HdDevG& d = dev_view(dv);
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        if (adv) advance_time(d);
        roll_repairs(d, 0);
    }
    int i, j;
    if (!row_cell(d, tiled, i, j)) return;
    const int N = d.N, M = d.M, W = d.W;
    const int p = i * W + j;
    const double sm = smooth_u_at(d, i, j);
    /* the row's wetness at j, j+1, M-2 and M-1: solveV's cells and the
     * runs' margins */
    const int* kh = d.khouot;
    const bool rowok = i >= 2 && i <= N;
    bool wx = false, wn = false, wl2 = false, wl1 = false;
    if (rowok) {
        wx = j >= 2 && j <= M && kh[p] == 0;
        wn = j + 1 >= 2 && j + 1 <= M && kh[p + 1] == 0;
        wl2 = kh[i * W + M - 2] == 0;
        wl1 = kh[i * W + M - 1] == 0;
    }
    const bool in = rowok && j >= 2 && j <= M - 1 && wx && wn && (j <= M - 2 || wl2);
    const DryMark dm = dry_load(d, i, j, 1);
    double vn = 0;
    if (in) {
        ExRaw r = v_solver_ld(d, i, j);
        r.t0 = smooth_u_at(d, i - 1, j);
        r.t1 = sm;
        r.t2 = smooth_u_at(d, i - 1, j + 1);
        r.t3 = smooth_u_at(d, i, j + 1);
        vn = v_solver_ev(d.P, r);
    }
    const EdgeVals ev = in ? EdgeVals{0.0, 0.0, 0.0, 0.0} : edge_vals<true>(d, i, j);
    d.tmp_u[p] = sm;
    if (in) d.t_v[p] = vn;
    dry_store(d, i, j, 1, dm);
    if (rowok) {
        int mp[3];
        double mv[3];
        const int nm = margin_at<true>(d, i, j, wx, wn, wl2, wl1, ev, mp, mv);
        for (int k = 0; k < nm; ++k) d.t_v[i * W + mp[k]] = mv[k];
    }
}

/* The VZ half-step's smoothing of t_v (Normalize + update_buffer into the
 * partner buffer, as k_normalize) and solveU (k_explicit_rows_u) in one pass
 * over the active row tiles: solveU reads the smoothed t_v at the four faces
 * around its cell, which this thread recomputes from the unsmoothed buffer
 * (the values the smoothing pass writes), so the two need no launch
 * boundary between them; the margins of the columns' runs (margin_at) and
 * the drying flags, wet bits and wetting candidates of the column reset
 * too. */
__global__ void __launch_bounds__(256) k_smooth_v_explicit_u(const HdDev* __restrict__ dv, int tiled, int adv)
{
    const HdDevG& d = dev_view(dv);
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        if (adv) advance_time(d);
        roll_repairs(d, 1);
    }
    int i, j;
    if (!row_cell(d, tiled, i, j)) return;
    const int N = d.N, M = d.M, W = d.W;
    const int p = i * W + j;
    const double sm = smooth_v_at(d, i, j);
    /* the column's wetness at i, i+1, N-2 and N-1: solveU's cells and the
     * runs' margins */
    const int* kh = d.khouot;
    const bool colok = j >= 2 && j <= M;
    bool wx = false, wn = false, wl2 = false, wl1 = false;
    if (colok) {
        wx = i >= 2 && i <= N && kh[p] == 0;
        wn = i + 1 >= 2 && i + 1 <= N && kh[p + W] == 0;
        wl2 = kh[(N - 2) * W + j] == 0;
        wl1 = kh[(N - 1) * W + j] == 0;
    }
    const bool in = colok && i >= 2 && i <= N - 1 && wx && wn && (i <= N - 2 || wl2);
    const DryMark dm = dry_load(d, i, j, 0);
    double un = 0;
    if (in) {
        ExRaw r = u_solver_ld(d, i, j);
        r.t0 = smooth_v_at(d, i, j - 1);
        r.t1 = sm;
        r.t2 = smooth_v_at(d, i + 1, j - 1);
        r.t3 = smooth_v_at(d, i + 1, j);
        un = u_solver_ev(d.P, r);
    }
    const EdgeVals ev = in ? EdgeVals{0.0, 0.0, 0.0, 0.0} : edge_vals<false>(d, j, i);
    d.tmp_v[p] = sm;
    if (in) d.t_u[p] = un;
    dry_store(d, i, j, 0, dm);
    if (colok) {
        int mp[3];
        double mv[3];
        const int nm = margin_at<false>(d, j, i, wx, wn, wl2, wl1, ev, mp, mv);
        for (int k = 0; k < nm; ++k) d.t_u[mp[k] * W + j] = mv[k];
    }
}

/* ================================================================ wet / dry */

/* Wetting pass of Reset_states_horizontal (support_funcs.cu:92-112) /
 * _vertical (:152-171), one wave per line.  The serial scan can only act at
 * positions next to a khouot == 1 cell; those were recorded by the drying
 * marks (dry_store, or k_normalize with post in the cross-check forms).
 * The wave sorts them (ranks by comparison), every lane loads the window of
 * one recorded position p at once (khouot and htaiz at p-1 .. p+kWetLA+1, z
 * at p .. p+kWetLA, t_z and FS at p), and lane 0 replays the scan over the
 * recorded positions in increasing order from LDS: a visit at q opens q-1
 * and q+1 where khouot == 1 and t_z(q) > z(q), copying q's t_z and FS;
 * opening q+1 makes q+1 the next visit, with q's values (the reference's
 * cascade).  Within a window every value is current except khouot at q-1
 * when an earlier visit opened it from the left, which is tracked; a
 * cascade past the window reads global memory.  With up to kWetFew
 * candidates lane 0 loads as it visits (one round of loads per visit); a
 * line with more candidates than the list holds is scanned in full. */
constexpr int kWetLA = 3;           /* cascade steps a window covers */
constexpr int kWetFew = 8;          /* up to this many candidates lane 0 loads as it visits */
constexpr int kWetWin = kWetLA + 3; /* khouot / htaiz entries p-1 .. p+kWetLA+1 */

__device__ __forceinline__ void wetting_line(const HdDevG& d, int horizontal, int line, int mode)
{
    const int lane = threadIdx.x;
    const int N = d.N, M = d.M, W = d.W;
    if (line < 1 || line > (horizontal ? N : M)) return;
    int* cand_cnt = d.cand_cnt;
    const int n = cand_cnt[line];
    if (n == 0) return;
    if (lane == 0) cand_cnt[line] = 0;
    const int pmax = horizontal ? M : N;
    double *tz = d.t_z, *Hm = d.H_moi, *FS = d.FS;
    const double *z = d.z, *ht = d.htaiz;
    int* kh = d.khouot;
    unsigned char* dryflag = d.dryflag;
    auto at = [&](int k) -> int { return horizontal ? line * W + k : k * W + line; };
    /* open cell q from a source with level tzs and concentration fss (the
     * reference copies t_z and FS from the source, H_moi = htaiz + t_z) */
    auto open = [&](int q, double tzs, double fss, double htq) {
        const int pq = at(q);
        const double hm = htq + tzs;
        tz[pq] = tzs;
        Hm[pq] = hm;
        kh[pq] = 0;
        FS[pq] = fss;
        bool inrange = horizontal ? (q >= 2 && q <= M && line >= 1 && line <= N) : (q >= 2 && q <= N && line >= 1 && line <= M);
        const bool dr = inrange && (hm <= d.P.H_TINH);
        dryflag[pq] = dr;
        if (!dr) { /* the opened cell counts as wet for the segment finder */
            const int r = horizontal ? line : q, cc = horizontal ? q : line;
            atomicOr((unsigned long long*)d.wetb + (size_t)r * ((M + 3 + 63) >> 6) + (cc >> 6), 1ull << (cc & 63));
        }
    };
    /* mode (a cross-check of the three ways, same bits): 0 as described,
     * 1 windows for every line, 2 the whole-line scan for every line */
    if (n > kCand || mode == 2) { /* more candidates than recorded: the whole line, one visit per round of loads */
        if (lane != 0) return;
        bool cascade = false;
        double tzc = 0, fsc = 0;
        for (int pos = 2; pos <= pmax; ++pos) {
            const int pp = at(pos), pm = at(pos - 1), pn = at(pos + 1);
            const double zc = z[pp];
            const double tzv = cascade ? tzc : tz[pp];
            const double fsv = cascade ? fsc : FS[pp];
            const int km = kh[pm], kn = kh[pn];
            const double htm = ht[pm], htn = ht[pn];
            cascade = false;
            if (tzv > zc) {
                if (km == 1) open(pos - 1, tzv, fsv, htm);
                if (kn == 1) {
                    open(pos + 1, tzv, fsv, htn);
                    cascade = true;
                    tzc = tzv;
                    fsc = fsv;
                }
            }
        }
        return;
    }
    __shared__ int s_pos[kCand];
    __shared__ int s_kh[kCand][kWetWin];
    __shared__ double s_ht[kCand][kWetWin], s_z[kCand][kWetLA + 1], s_tz[kCand], s_fs[kCand];
    /* the recorded positions (distinct), sorted by rank */
    const int c = lane < n ? d.cand[line * kCand + lane] : 0x7fffffff;
    int rank = 0;
    for (int k = 0; k < n; ++k) rank += __shfl(c, k, kWave) < c;
    if (lane < n) s_pos[rank] = c;
    __syncthreads();
    const bool win = n > (mode == 1 ? 0 : kWetFew); /* with few candidates lane 0 loads as it visits */
    if (win && lane < n) { /* the window of one position, all loads at once */
        const int p = s_pos[lane];
        s_tz[lane] = tz[at(p)];
        s_fs[lane] = FS[at(p)];
#pragma unroll
        for (int o = 0; o < kWetWin; ++o) {
            const int q = p - 1 + o;
            const bool in = q <= pmax + 1;
            s_kh[lane][o] = in ? kh[at(q)] : 2;
            s_ht[lane][o] = in ? ht[at(q)] : 0.0;
        }
#pragma unroll
        for (int o = 0; o <= kWetLA; ++o) s_z[lane][o] = p + o <= pmax ? z[at(p + o)] : 0.0;
    }
    __syncthreads();
    if (lane != 0) return;
    /* the scan: recorded positions in order, cascades after right openings */
    int li = 0, pos = 1, bw = 0;
    int lastR = -8, prevR = -8; /* the last two cells opened from the left neighbour */
    bool cascade = false;
    double tzc = 0, fsc = 0;
    while (true) {
        if (cascade) {
            ++pos;
        } else {
            while (li < n && s_pos[li] <= pos) ++li;
            if (li >= n) break;
            pos = s_pos[li];
            bw = li;
        }
        if (pos < 2 || pos > pmax) break;
        const int o = pos - s_pos[bw];
        double zc, htm, htn, tzv, fsv;
        int km, kn;
        if (win && o <= kWetLA) {
            zc = s_z[bw][o];
            km = s_kh[bw][o];
            kn = s_kh[bw][o + 2];
            htm = s_ht[bw][o];
            htn = s_ht[bw][o + 2];
            tzv = cascade ? tzc : s_tz[bw];
            fsv = cascade ? fsc : s_fs[bw];
            if (pos - 1 == lastR || pos - 1 == prevR) km = 0;
        } else { /* current memory (this lane's own stores included) */
            const int pp = at(pos), pm = at(pos - 1), pn = at(pos + 1);
            zc = z[pp];
            km = kh[pm];
            kn = kh[pn];
            htm = ht[pm];
            htn = ht[pn];
            tzv = cascade ? tzc : tz[pp];
            fsv = cascade ? fsc : FS[pp];
        }
        cascade = false;
        if (tzv > zc) {
            if (km == 1) open(pos - 1, tzv, fsv, htm);
            if (kn == 1) {
                open(pos + 1, tzv, fsv, htn);
                prevR = lastR;
                lastR = pos + 1;
                cascade = true;
                tzc = tzv;
                fsc = fsv;
            }
        }
    }
}

__global__ void __launch_bounds__(64) k_wetting(const HdDev* __restrict__ dv, int horizontal, int mode)
{
    const HdDevG& d = dev_view(dv);
    /* the run tables are consumed for this half-step (solver and sediment
     * sweeps); the segment finder at its end appends from zero */
    if (blockIdx.x == 0 && threadIdx.x == 0) { d.run_cnt[0] = 0; d.run_cnt[1] = 0; }
    wetting_line(d, horizontal, blockIdx.x, mode);
}

/* The second smoothing pass of a half-step (Normalize + update_buffer of the
 * velocity the explicit pass produced, support_funcs.cu:682-729, into the
 * partner buffer) with update_h_moi (support_funcs.cu:60-73), and the
 * wetting pass of Reset_states, in one launch: workgroups [0, nlb) are the
 * wetting lines (wetting_line), the others smooth one 64-cell tile row each
 * (or every row of the grid when not tiled).  The smoothing weighs the
 * neighbours across by the mask before the wetting (the kh0b bits the
 * explicit pass wrote), and update_h_moi covers the cells wet before the
 * wetting (which writes H_moi of the cells it opens itself), so neither half
 * reads what the other writes. */
__global__ void __launch_bounds__(64) k_smooth_wet(const HdDev* __restrict__ dv, int isU, int horizontal, int tiled,
                                                    int nlb, int mode)
{
    const HdDevG& d = dev_view(dv);
    if ((int)blockIdx.x < nlb) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { d.run_cnt[0] = 0; d.run_cnt[1] = 0; }
        wetting_line(d, horizontal, blockIdx.x, mode);
        return;
    }
    const int t = (int)blockIdx.x - nlb;
    const int ncb = (d.M + 3 + 63) >> 6;
    int tile;
    if (tiled) {
        if (t >= d.ntiles[0]) return;
        tile = d.tiles[t];
    } else {
        if (t >= (d.N + 3) * ncb) return;
        tile = t;
    }
    const int i = tile / ncb, j = (tile % ncb) * 64 + (int)threadIdx.x;
    const int N = d.N, M = d.M, W = d.W;
    if (j >= M + 3) return;
    const int p = i * W + j;
    const double* val = isU ? d.t_u : d.t_v;
    double* out = isU ? d.tmp_u : d.tmp_v;
    const int close = isU ? W : 1;
    const bool in = i >= 2 && i <= N && j >= 2 && j <= M;
    /* the neighbours across: (i, j -+ 1) for u, (i -+ 1, j) for v */
    const int fim = isU ? i : i - 1, fjm = isU ? j - 1 : j, fip = isU ? i : i + 1, fjp = isU ? j + 1 : j;
    auto kh0 = [&](int qi, int qj) -> bool {
        return (d.kh0b[(size_t)qi * ncb + (qj >> 6)] >> (qj & 63)) & 1;
    };
    const double r0 = val[p];
    const double vcm = val[in ? p - close : p], vcp = val[in ? p + close : p];
    const double vfm = val[in ? fim * W + fjm : p], vfp = val[in ? fip * W + fjp : p];
    const bool kfm = in ? kh0(fim, fjm) : false, kfp = in ? kh0(fip, fjp) : false;
    const bool wet0 = kh0(i, j);
    const double htz = d.htaiz[p], tz = d.t_z[p];
    double r = r0;
    if (in && r != 0) {
        int count = 2;
        double nb = vcm + vcp;
        if (kfm) { nb += vfm; count++; }
        if (kfp) { nb += vfp; count++; }
        r = r * d.P.heso + (1 - d.P.heso) * nb / count;
    }
    out[p] = r;
    if (i >= 1 && i <= N && j >= 1 && j <= M && wet0) d.H_moi[p] = htz + tz;
}

/* update_uvz (support_funcs.cu:671-680) fused with the drying writes of
 * Reset_states (support_funcs.cu:114-127 / 173-185) and Htuongdoi
 * (support_funcs.cu:338-354).  Drying cell (a,b) zeroes t_u at (a-1,b),(a,b)
 * and t_v at (a,b-1),(a,b); each cell applies the flags of itself and of the
 * neighbours that write it, then copies.  Face depths read the new z (t_z)
 * inside [0,N]x[0,M] and the untouched z outside, as after update_uvz. */
__device__ __forceinline__ void uvz_cell(const HdDevG& d, int i, int j, int kd, int kn)
{
    const unsigned char* dryflag = d.dryflag;
    int N = d.N, M = d.M, W = d.W;
    int p = i * W + j;
    bool me = dryflag[p];
    bool dn = (i + 1 <= N + 2) ? dryflag[p + W] : false;
    bool rt = (j + 1 <= M + 2) ? dryflag[p + 1] : false;
    double tu = d.t_u[p], tv = d.t_v[p];
    if (me || dn) { tu = 0; d.t_u[p] = 0; }
    if (me || rt) { tv = 0; d.t_v[p] = 0; }
    if (me) { d.khouot[p] = 1; d.FS[p] = 0; }
    if (i <= N && j <= M) {
        d.z[p] = d.t_z[p];
        d.u[p] = tu * (1 - kd);
        d.v[p] = tv * (1 - kn);
    }
    if (i >= 1 && i <= N && j >= 1 && j <= M) {
        const double* h = d.h;
        double zc = d.t_z[p];
        double zd = (i + 1 <= N) ? d.t_z[p + W] : d.z[p + W];
        double zr = (j + 1 <= M) ? d.t_z[p + 1] : d.z[p + 1];
        d.Htdu[p] = (h[i * W + j - 1] + h[p] + zd + zc) * 0.5;
        d.Htdv[p] = (h[(i - 1) * W + j] + h[p] + zr + zc) * 0.5;
    }
}

__global__ void __launch_bounds__(256) k_uvz(const HdDev* __restrict__ dv, int kd, int kn, int tiled)
{
    const HdDevG& d = dev_view(dv);
    /* the run tables are rebuilt by the next kernel on the stream */
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) { d.run_cnt[0] = 0; d.run_cnt[1] = 0; }
    int i, j;
    if (!row_cell(d, tiled, i, j)) return;
    uvz_cell(d, i, j, kd, kn);
}

/* After u, v, z were written from outside the step: t_u = u, t_v = v,
 * t_z = z everywhere; H_moi = htaiz + z on wet cells of [1,N]x[1,M]
 * (update_h_moi, support_funcs.cu:60-73); face depths from z (Htuongdoi,
 * support_funcs.cu:338-354). */
__global__ void __launch_bounds__(256) k_flow_refresh(const HdDev* __restrict__ dv)
{
    const HdDevG& d = dev_view(dv);
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    int N = d.N, M = d.M, W = d.W;
    if (j >= M + 3 || i >= N + 3) return;
    int p = i * W + j;
    d.t_u[p] = d.u[p];
    d.t_v[p] = d.v[p];
    d.t_z[p] = d.z[p];
    if (i >= 1 && i <= N && j >= 1 && j <= M) {
        if (d.khouot[p] == 0) d.H_moi[p] = d.htaiz[p] + d.z[p];
        const double* h = d.h;
        const double* z = d.z;
        d.Htdu[p] = (h[i * W + j - 1] + h[p] + z[p + W] + z[p]) * 0.5;
        d.Htdv[p] = (h[(i - 1) * W + j] + h[p] + z[p + 1] + z[p]) * 0.5;
    }
}

/* dH += f * (dH - snap) over the padded grid (bed-change extrapolation) */
__global__ void __launch_bounds__(256) k_bed_extrapolate(const HdDev* __restrict__ dv, const double* snap, double f)
{
    const HdDevG& d = dev_view(dv);
    const long n = (long)(d.N + 3) * d.W;
    for (long p = blockIdx.x * 256L + threadIdx.x; p < n; p += gridDim.x * 256L) {
        double x = d.dH[p];
        d.dH[p] = x + f * (x - snap[p]);
    }
}

/* Over the wet cells (khouot == 0) of [1,N]x[1,M]: max |z - zref| into
 * out[0] (as the bit pattern of a non-negative double, compared as an
 * unsigned integer), sum of htaiz + z into out[1], sum of htaiz + zref into
 * out[2] (FP64 atomics; the sums' last bits depend on the order), the number
 * of cells with |z - zref| > bound (or NaN) into out[3] and the number of wet
 * cells into out[4] (integers below 2^53, exact in FP64). */
__global__ void __launch_bounds__(256) k_z_change(const HdDev* __restrict__ dv, const double* zref, double bound,
                                                  double* out)
{
    const HdDevG& d = dev_view(dv);
    const int N = d.N, M = d.M, W = d.W;
    __shared__ double s1[256], s2[256], s3[256], s4[256];
    __shared__ unsigned long long smx[256];
    double v1 = 0, v2 = 0, nx = 0, nw = 0;
    unsigned long long mx = 0;
    const long n = (long)(N + 3) * W;
    for (long p = blockIdx.x * 256L + threadIdx.x; p < n; p += gridDim.x * 256L) {
        int i = (int)(p / W), j = (int)(p % W);
        if (i < 1 || i > N || j < 1 || j > M || d.khouot[p] != 0) continue;
        double dz = fabs(d.z[p] - zref[p]);
        unsigned long long b = __double_as_longlong(dz);
        if (!(dz <= 1e300)) b = 0x7FF0000000000000ull; /* NaN counts as infinite */
        if (b > mx) mx = b;
        if (!(dz <= bound)) nx += 1;
        nw += 1;
        v1 += d.htaiz[p] + d.z[p];
        v2 += d.htaiz[p] + zref[p];
    }
    s1[threadIdx.x] = v1;
    s2[threadIdx.x] = v2;
    s3[threadIdx.x] = nx;
    s4[threadIdx.x] = nw;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            s1[threadIdx.x] += s1[threadIdx.x + k];
            s2[threadIdx.x] += s2[threadIdx.x + k];
            s3[threadIdx.x] += s3[threadIdx.x + k];
            s4[threadIdx.x] += s4[threadIdx.x + k];
            if (smx[threadIdx.x + k] > smx[threadIdx.x]) smx[threadIdx.x] = smx[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMax((unsigned long long*)out, smx[0]);
        atomicAdd(out + 1, s1[0]);
        atomicAdd(out + 2, s2[0]);
        atomicAdd(out + 3, s3[0]);
        atomicAdd(out + 4, s4[0]);
    }
}

/* ================================================================ segments */

/* Interpolate_FS_ng (support_funcs.cu:189-208) of row N (p = 0) or row 2
 * (p = 1) at column j, and Interpolate_FS_d (support_funcs.cu:210-229) of
 * column M at row i: whether the cell moves (t > 0), its travel distance and
 * its source row / column (clamped to [0, N+1] / [0, M+1]).  `dry` (the
 * drying flags, when the state copy of the half-step runs in the same
 * launch) gives the velocities that copy leaves: t zeroed at drying cells,
 * u = t_u (1 - kd), v = t_v (1 - kn) (rows 0..N). */
struct FsEdge {
    bool on;
    double dl;
    int src;
};

__device__ __forceinline__ FsEdge fs_row_edge(const HdDevG& d, const unsigned char* dry, int kd, int p, int j)
{
    const int N = d.N, W = d.W, loc = p == 0 ? N : 2;
    FsEdge e{false, 0.0, 0};
    if (j < 1 || j >= d.M) return e;
    double tq = d.t_u[loc * W + j], q = d.u[loc * W + j];
    if (dry) {
        if (dry[loc * W + j] || dry[(loc + 1) * W + j]) tq = 0;
        q = tq * (1 - kd);
    }
    e.on = tq > 0;
    e.dl = (q + tq) * 0.25 * d.P.dT;
    const int offset = p == 0 ? N - 1 : 2, sign = p == 0 ? -1 : 1;
    int r = (int)(offset + sign * (e.dl / d.P.dX));
    if (r < 0) r = 0;
    if (r > N + 1) r = N + 1;
    e.src = r;
    return e;
}

__device__ __forceinline__ FsEdge fs_col_edge(const HdDevG& d, const unsigned char* dry, int kn, int i)
{
    const int N = d.N, M = d.M, W = d.W;
    FsEdge e{false, 0.0, 0};
    double tq = d.t_v[i * W + M], q = d.v[i * W + M];
    if (dry) {
        if (dry[i * W + M] || dry[i * W + M + 1]) tq = 0;
        if (i <= N) q = tq * (1 - kn);
    }
    e.on = tq > 0;
    e.dl = (q + tq) * 0.25 * d.P.dT;
    int r = (int)((M - 1) + -1 * (e.dl / d.P.dY));
    if (r < 0) r = 0;
    if (r > M + 1) r = M + 1;
    e.src = r;
    return e;
}

/* FS as the interpolations read it: zero where the state copy of the same
 * launch dries the cell */
__device__ __forceinline__ double fs_src(const HdDevG& d, const unsigned char* dry, int q)
{
    return (dry && dry[q]) ? 0.0 : d.FS[q];
}

__device__ __forceinline__ double fs_row_value(const HdDevG& d, int p, const FsEdge& e, double f0, double f1)
{
    const int i_dX = hd_rint_int(d.P.dX);
    const int q = i_dX != 0 ? hd_rint_int(e.dl) / i_dX : 0;
    const int sign = p == 0 ? -1 : 1;
    return f0 + sign * (f1 - f0) * (e.dl / d.P.dX - q);
}

__device__ __forceinline__ double fs_col_value(const HdDevG& d, const FsEdge& e, double f0, double f1)
{
    const int j_dY = hd_rint_int(d.P.dY);
    const int q = j_dY != 0 ? hd_rint_int(e.dl) / j_dY : 0;
    return f1 + (f1 - f0) * (e.dl / d.P.dY - q);
}

/* The new values of rows N and 2 at column j (the reference runs row N,
 * then row 2, then column M).  Each reads only its own column, so row 2
 * sees row N's new value at (N, j) from the same thread, and row N reads
 * (2, j) before this thread writes it. */
__device__ __forceinline__ void fs_rows_at(const HdDevG& d, const unsigned char* dry, int kd, int j, bool store,
                                           bool& onN, double& vN, bool& on2, double& v2)
{
    const int N = d.N, W = d.W;
    const FsEdge e1 = fs_row_edge(d, dry, kd, 0, j), e2 = fs_row_edge(d, dry, kd, 1, j);
    double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
    if (e1.on) { a0 = fs_src(d, dry, e1.src * W + j); a1 = fs_src(d, dry, (e1.src + 1) * W + j); }
    if (e2.on) { b0 = fs_src(d, dry, e2.src * W + j); b1 = fs_src(d, dry, (e2.src + 1) * W + j); }
    onN = e1.on;
    on2 = e2.on;
    if (e1.on) {
        vN = fs_row_value(d, 0, e1, a0, a1);
        if (store) d.FS[N * W + j] = vN;
        if (e2.src == N) b0 = vN;
        if (e2.src + 1 == N) b1 = vN;
    }
    if (e2.on) {
        v2 = fs_row_value(d, 1, e2, b0, b1);
        if (store) d.FS[2 * W + j] = v2;
    }
}

/* The FS edge interpolations, spread over workgroups of 256 without any
 * ordering between them: workgroup k < nfsr takes columns 1 + 256k + t of
 * rows N and 2; the others take rows i = 1 + 256 (k - nfsr) + t of column M.
 * The column cell of row N (row 2) reads that row's new values at its two
 * source columns, and computes them itself. */
__device__ void fs_edges_block(const HdDevG& d, int k, int nfsr, const unsigned char* dry, int kd, int kn)
{
    const int N = d.N, M = d.M, W = d.W;
    if (k < nfsr) {
        const int j = 1 + 256 * k + (int)threadIdx.x;
        bool onN, on2;
        double vN, v2;
        fs_rows_at(d, dry, kd, j, true, onN, vN, on2, v2);
        return;
    }
    const int imax = (M - 1 < N + 2) ? M - 1 : N + 2;
    const int i = 1 + 256 * (k - nfsr) + (int)threadIdx.x;
    if (i > imax) return;
    const FsEdge e = fs_col_edge(d, dry, kn, i);
    if (!e.on) return;
    double f0 = fs_src(d, dry, i * W + e.src), f1 = fs_src(d, dry, i * W + e.src + 1);
    if (i == N || i == 2) {
        bool onN, on2;
        double vN, v2;
        fs_rows_at(d, dry, kd, e.src, false, onN, vN, on2, v2);
        if (i == N ? onN : on2) f0 = i == N ? vN : v2;
        fs_rows_at(d, dry, kd, e.src + 1, false, onN, vN, on2, v2);
        if (i == N ? onN : on2) f1 = i == N ? vN : v2;
    }
    d.FS[i * W + M] = fs_col_value(d, e, f0, f1);
}

__global__ void k_reset_runs(const HdDev* __restrict__ dv)
{
    const HdDevG& d = dev_view(dv);
    if (threadIdx.x == 0) { d.run_cnt[0] = 0; d.run_cnt[1] = 0; }
}

/* Wet bits: wetb[i * ncb + w] (ncb = (M + 3 + 63) / 64), bit b, is set
 * when cell (i, 64 w + b) has khouot == 0 and no drying flag, i.e. the cells
 * the segment finder counts as wet.  The drying pass (k_normalize with post)
 * writes every word of the cells it covers, one ballot per 64-cell tile row,
 * and the wetting pass sets the bits of the cells it opens; k_wet_bits
 * builds them from khouot alone after init and restarts.  Words of tiles
 * that hold only khouot == 2 cells stay zero. */

/* bits 0 .. n of a word (none for n < 0, all for n >= 63) */
__device__ __forceinline__ unsigned long long bits_upto(int n)
{
    return n < 0 ? 0ull : (n >= 63 ? ~0ull : ((2ull << n) - 1));
}

/* Starts and ends of the wet runs among positions base .. base + 63 of a
 * line (the rule of the serial scan, support_funcs.cu:252-279: a run starts
 * at a wet position <= L-2 whose predecessor is dry; it ends at its last wet
 * position, where an end at L-1 counts only when L-2 is wet and an end at L
 * only when L-2 and L-1 are).  m: the positions' wet bits, already limited
 * to 2 .. L; prev / next: wetness of base - 1 and base + 64. */
struct RunBits {
    unsigned long long st, en;
};
__device__ __forceinline__ RunBits run_bits(unsigned long long m, int base, int L, unsigned prev, unsigned next,
                                            unsigned wl2, unsigned wl1)
{
    const unsigned long long le2 = bits_upto(L - 2 - base);
    unsigned long long extra = 0;
    if (wl2 && L - 1 >= base && L - 1 < base + 64) extra |= 1ull << (L - 1 - base);
    if (wl2 && wl1 && L >= base && L < base + 64) extra |= 1ull << (L - base);
    RunBits r;
    r.st = m & ~((m << 1) | (unsigned long long)prev) & le2;
    r.en = m & ~((m >> 1) | ((unsigned long long)next << 63)) & (le2 | extra);
    return r;
}

/* The runs of row i (positions 2..M) by one wave from its wet bits: a word
 * per lane (64 words = 4096 cells per batch), neighbours' edge bits by
 * shuffles, ranks by a wave prefix sum.  Entries k < SL go to dau / cuoi
 * (global) and sf / sl (LDS); returns the number of runs found. */
__device__ __forceinline__ int row_word_scan(const HdDevG& d, int i, int lane, int* dau, int* cuoi, int* sf, int* sl)
{
    const int M = d.M, L = M, SL = d.SL;
    const int ncb = (M + 3 + 63) >> 6;
    const auto row = d.wetb + (size_t)i * ncb;
    auto wetx = [&](int x) -> unsigned {
        return (x >= 2 && x <= L) ? (unsigned)((row[x >> 6] >> (x & 63)) & 1) : 0u;
    };
    const unsigned wl2 = wetx(L - 2), wl1 = wetx(L - 1);
    unsigned carry = 0; /* wetness of the position before the batch */
    int ns = 0, ne = 0;
    for (int w0 = 0; w0 < ncb; w0 += kWave) {
        const int w = w0 + lane, base = 64 * w;
        unsigned long long m = w < ncb ? row[w] : 0ull;
        m &= bits_upto(L - base) & ~bits_upto(1 - base);
        const unsigned after = wetx(64 * (w0 + kWave));
        const unsigned long long pw = __shfl_up(m, 1, kWave), nw = __shfl_down(m, 1, kWave);
        const unsigned prev = lane == 0 ? carry : (unsigned)(pw >> 63);
        const unsigned next = lane == kWave - 1 ? after : (unsigned)(nw & 1);
        carry = (unsigned)(__shfl(m, kWave - 1, kWave) >> 63);
        const RunBits rb = run_bits(m, base, L, prev, next, wl2, wl1);
        const int cs = __popcll(rb.st), ce = __popcll(rb.en);
        int ps = cs, pe = ce; /* inclusive prefix over the lanes */
        for (int o = 1; o < kWave; o <<= 1) {
            const int a = __shfl_up(ps, o, kWave), b = __shfl_up(pe, o, kWave);
            if (lane >= o) { ps += a; pe += b; }
        }
        int ks = ns + ps - cs, ke = ne + pe - ce;
        for (unsigned long long b = rb.st; b; b &= b - 1, ++ks) {
            const int x = base + __ffsll(b) - 1;
            if (ks < SL) { dau[ks] = x; sf[ks] = x; }
        }
        for (unsigned long long b = rb.en; b; b &= b - 1, ++ke) {
            const int x = base + __ffsll(b) - 1;
            if (ke < SL) { cuoi[ke] = x; sl[ke] = x; }
        }
        ns += __shfl(ps, kWave - 1, kWave);
        ne += __shfl(pe, kWave - 1, kWave);
    }
    return ns;
}

/* Transpose of a 64 x 64 bit block held one row per lane (bit q of lane r
 * = element (r, q)): afterwards lane q holds column q (bit r = element
 * (r, q)).  Six rounds; in round s the lanes r and r ^ s swap the
 * off-diagonal s x s sub-blocks of their 2s x 2s block. */
__device__ __forceinline__ unsigned long long transpose64(unsigned long long a, int lane)
{
    constexpr unsigned long long kKeep[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                                             0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for (int k = 0, sh = 32; k < 6; ++k, sh >>= 1) {
        const unsigned long long b = __shfl_xor(a, sh, kWave);
        if (lane & sh) a ^= ((b >> sh) ^ a) & kKeep[k];
        else a ^= (((a >> sh) ^ b) & kKeep[k]) << sh;
    }
    return a;
}

/* The column tables of Find_Calculation_limits_Vertical (the run rule of
 * run_bits, support_funcs.cu:303-330) for the 64 columns of word column cb,
 * by a block of four waves: in a pass of 1024 rows, lane r of wave c loads
 * the words of rows 64 (4c + k) + r (k < 4; one word carries the 64
 * columns), four transposes give each lane its column's bits of those rows,
 * the waves' edge bits and run counts meet in LDS, ranks are prefix sums
 * over the waves.  Rows beyond 1024 are done in further passes with the
 * counts and the last bit carried.  Every thread returns its column's
 * number of entries (min(runs, SL)); entries go to dauj / cuoij and to
 * sf / sl [column * 24 + k], and s_cbase[column] receives the column's
 * first slot in the packed run table (reserved in the last pass).  Ends
 * with a barrier. */
constexpr int kColKW = 4; /* words (64 rows each) per wave and pass */

__device__ __forceinline__ int col_word_scan(const HdDevG& d, int cb, int* s_f, int* s_l, int* s_cbase)
{
    const int N = d.N, M = d.M, SL = d.SL, L = N;
    const int ncb = (M + 3 + 63) >> 6;
    const int lane = threadIdx.x & (kWave - 1), c = threadIdx.x / kWave;
    const int j = 64 * cb + lane;
    const bool colok = j >= 2 && j <= M;
    const auto wb = d.wetb + cb; /* word (x, cb) at wb[x * ncb] */
    auto wetx = [&](int x) -> unsigned {
        return (colok && x >= 2 && x <= L) ? (unsigned)((wb[(size_t)x * ncb] >> lane) & 1) : 0u;
    };
    __shared__ unsigned s_first[4][kWave], s_last[4][kWave], s_pb[kWave];
    __shared__ int s_cs[4][kWave], s_ce[4][kWave], s_ts[kWave], s_te[kWave];
    if (c == 0) { s_ts[lane] = 0; s_te[lane] = 0; s_pb[lane] = 0; }
    __syncthreads();
    /* after the barrier, so that these loads travel with the first pass's */
    const unsigned wl2 = wetx(L - 2), wl1 = wetx(L - 1);
    constexpr int kPass = 4 * kColKW * kWave;
    for (int x0 = 0; x0 <= L; x0 += kPass) {
        unsigned long long t[kColKW], st[kColKW], en[kColKW];
#pragma unroll
        for (int k = 0; k < kColKW; ++k) {
            const int r = x0 + kWave * (kColKW * c + k) + lane;
            t[k] = r <= L ? wb[(size_t)r * ncb] : 0ull;
        }
        const unsigned nxt = wetx(x0 + kPass);
#pragma unroll
        for (int k = 0; k < kColKW; ++k) {
            const int base = x0 + kWave * (kColKW * c + k);
            t[k] = transpose64(t[k], lane) & (colok ? (bits_upto(L - base) & ~bits_upto(1 - base)) : 0ull);
        }
        s_first[c][lane] = (unsigned)(t[0] & 1);
        s_last[c][lane] = (unsigned)(t[kColKW - 1] >> 63);
        __syncthreads();
        unsigned prev = c > 0 ? s_last[c - 1][lane] : s_pb[lane];
        const unsigned after = c < 3 ? s_first[c + 1][lane] : nxt;
        int cs = 0, ce = 0;
#pragma unroll
        for (int k = 0; k < kColKW; ++k) {
            const int base = x0 + kWave * (kColKW * c + k);
            const unsigned nk = k < kColKW - 1 ? (unsigned)(t[k + 1] & 1) : after;
            const RunBits rb = run_bits(t[k], base, L, prev, nk, wl2, wl1);
            prev = (unsigned)(t[k] >> 63);
            st[k] = rb.st;
            en[k] = rb.en;
            cs += __popcll(rb.st);
            ce += __popcll(rb.en);
        }
        s_cs[c][lane] = cs;
        s_ce[c][lane] = ce;
        __syncthreads();
        if (c == 0 && x0 + kPass > L) {
            /* last pass: the columns' entry counts are known; reserve their
             * slots of the packed run table (column order, one atomic) while
             * the entries are written */
            int ns = s_ts[lane];
            for (int w = 0; w < 4; ++w) ns += s_cs[w][lane];
            const int nn = colok ? (ns < SL ? ns : SL) : 0;
            int p = nn;
            for (int o = 1; o < kWave; o <<= 1) {
                const int a = __shfl_up(p, o, kWave);
                if (lane >= o) p += a;
            }
            const int tot = __shfl(p, kWave - 1, kWave);
            int base = 0;
            if (lane == 0 && tot) base = atomicAdd(d.run_cnt + 0, tot);
            s_cbase[lane] = __shfl(base, 0, kWave) + p - nn;
        }
        int ks = s_ts[lane], ke = s_te[lane];
        for (int w = 0; w < c; ++w) { ks += s_cs[w][lane]; ke += s_ce[w][lane]; }
#pragma unroll
        for (int k = 0; k < kColKW; ++k) {
            const int base = x0 + kWave * (kColKW * c + k);
            for (unsigned long long b = st[k]; b; b &= b - 1, ++ks) {
                const int x = base + __ffsll(b) - 1;
                if (ks < SL) { d.dauj[j * SL + ks] = x; s_f[lane * 24 + ks] = x; }
            }
            for (unsigned long long b = en[k]; b; b &= b - 1, ++ke) {
                const int x = base + __ffsll(b) - 1;
                if (ke < SL) { d.cuoij[j * SL + ke] = x; s_l[lane * 24 + ke] = x; }
            }
        }
        __syncthreads(); /* the pass's counts are read before they are replaced */
        if (c == 3) {
            s_ts[lane] = ks;
            s_te[lane] = ke;
            s_pb[lane] = (unsigned)(t[kColKW - 1] >> 63);
        }
        __syncthreads();
    }
    const int ns = s_ts[lane];
    const int n = colok ? (ns < SL ? ns : SL) : 0;
    if (c == 0 && colok) {
        d.mocj[j] = n;
        if (ns > SL) d.flags[0] = 1;
    }
    return n;
}

/* Find_Calculation_limits_Horizontal / _Vertical (support_funcs.cu:230-335)
 * from the wet bits: the per-line tables daui / cuoii / moci (rows
 * i = 2..N) and dauj / cuoij / mocj (columns j = 2..M), plus the packed run
 * tables the solvers read (runs_v for rows, runs_u for columns; their
 * counters were zeroed earlier in the half-step).  Blocks of 256 threads;
 * each reserves its table slots with one atomic.  Blocks [0, nrowb): rows
 * 4b .. 4b+3, one wave per row; with count_wet these waves also add the
 * row's wet cells (rows 0..N+2) to the wet-cell accumulator.  Blocks
 * [nrowb, nrowb + ncolb): the 64 columns of word column b - nrowb
 * (col_word_scan).  The next nfs blocks: the FS edge interpolations
 * (fs_edges_block, nfsr of them for the rows); `dry` (the drying flags, when
 * the state copy runs in the same launch) marks the cells that copy
 * zeroes. */
__device__ __forceinline__ void find_limits_block(const HdDevG& d, int bid, int nrowb, int ncolb, int nfsr,
                                                  int count_wet, const unsigned char* dry, int kd, int kn)
{
    if (bid >= nrowb + ncolb) { /* the FS edge interpolations (independent of the mask) */
        fs_edges_block(d, bid - nrowb - ncolb, nfsr, dry, kd, kn);
        return;
    }
    const int N = d.N, M = d.M, SL = d.SL;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int ncb = (M + 3 + 63) >> 6;
    if (bid >= nrowb) { /* 64 columns per block */
        __shared__ int s_cf[kWave * 24], s_cl[kWave * 24], s_cbase[kWave];
        const int cb = bid - nrowb;
        if (64 * cb > M) return; /* no column of this context (a smaller member of a batch) */
        const int n = col_word_scan(d, cb, s_cf, s_cl, s_cbase); /* ends with a barrier */
        const int cap = (M + 3) * SL;
        const int j = 64 * cb + lane;
        for (int k = wv; k < n; k += 4) {
            const int at = s_cbase[lane] + k;
            if (at < cap) d.runs_u[at] = make_int4(j, s_cf[lane * 24 + k], s_cl[lane * 24 + k], 0);
        }
        return;
    }
    __shared__ int s_f[4][24], s_l[4][24], s_n[4], s_base[4];
    int n = 0;
    const int i = 4 * bid + wv; /* rows 0..N+2 for counting, 2..N for runs */
    if (count_wet && i <= N + 2) {
        int c = 0;
        for (int w = lane; w < ncb; w += kWave) c += __popcll(d.wetb[(size_t)i * ncb + w]);
        for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, kWave);
        if (lane == 0 && c) atomicAdd((unsigned long long*)d.wet_count, (unsigned long long)c);
    }
    if (i >= 2 && i <= N) {
        const int ns = row_word_scan(d, i, lane, d.daui + i * SL, d.cuoii + i * SL, s_f[wv], s_l[wv]);
        n = ns < SL ? ns : SL;
        if (lane == 0) {
            d.moci[i] = n;
            if (ns > SL) d.flags[0] = 1;
        }
    }
    if (lane == 0) s_n[wv] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = s_n[0] + s_n[1] + s_n[2] + s_n[3];
        int base = tot ? atomicAdd(d.run_cnt + 1, tot) : 0;
        for (int w = 0; w < 4; ++w) { s_base[w] = base; base += s_n[w]; }
    }
    __syncthreads();
    const int cap = (N + 3) * SL;
    const int at = s_base[wv] + lane;
    if (lane < n && at < cap) d.runs_v[at] = make_int4(i, s_f[wv][lane], s_l[wv][lane], 0);
}

/* The wet bits from khouot alone (no drying flags pending), for the segment
 * finder after init and restarts: one ballot per 64-cell tile row. */
__global__ void __launch_bounds__(256) k_wet_bits(const HdDev* __restrict__ dv)
{
    const HdDevG& d = dev_view(dv);
    const int ncb = (d.M + 3 + 63) >> 6;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6), i = blockIdx.y;
    if (w >= ncb || i >= d.N + 3) return; /* whole waves */
    const int j = 64 * w + (threadIdx.x & 63);
    const bool wet = j < d.M + 3 && d.khouot[(size_t)i * d.W + j] == 0;
    const unsigned long long bits = __ballot(wet);
    if ((threadIdx.x & 63) == 0) d.wetb[(size_t)i * ncb + w] = bits;
}

__global__ void __launch_bounds__(256) k_find_limits(const HdDev* __restrict__ dv, int nrowb, int ncolb, int nfsr,
                                                      int count_wet)
{
    find_limits_block(dev_view(dv), blockIdx.x, nrowb, ncolb, nfsr, count_wet, nullptr, 0, 0);
}

/* The state copy (k_uvz) and the segment tables + FS edges (k_find_limits)
 * of one half-step in one launch: blocks [0, nseg) build the tables and
 * interpolate the FS edges, the rest copy the state over the active row
 * tiles (or the whole grid).  The table blocks do not wait for the copy:
 * the wet bits already exclude the cells it dries (it sets khouot = 1
 * exactly at flagged cells), and the FS edge blocks apply the copy's
 * zeroing and u = t_u (1 - kd), v = t_v (1 - kn) to what they read, so
 * either order of the two halves gives the result of copying first.  The
 * run counters were cleared by the wetting kernel. */
__global__ void __launch_bounds__(256) k_uvz_segments(const HdDev* __restrict__ dv, int kd, int kn, int tiled,
                                                       int nrowb, int ncolb, int nfsr, int nfs, int count_wet)
{
    const HdDevG& d = dev_view(dv);
    const int nseg = nrowb + ncolb + nfs;
    if ((int)blockIdx.x < nseg) {
        /* the column scans, the launch's longest workgroups, are dispatched
         * first: workgroup b < ncolb takes column block b */
        const int b = blockIdx.x;
        const int bid = b < ncolb ? nrowb + b : (b < ncolb + nrowb ? b - ncolb : b);
        find_limits_block(d, bid, nrowb, ncolb, nfsr, count_wet, d.dryflag, kd, kn);
        return;
    }
    const int t = ((int)blockIdx.x - nseg) * 4 + (threadIdx.x >> 6);
    const int ncb = (d.M + 3 + 63) >> 6;
    int tile;
    if (tiled) {
        if (t >= d.ntiles[0]) return;
        tile = d.tiles[t];
    } else {
        if (t >= (d.N + 3) * ncb) return;
        tile = t;
    }
    const int i = tile / ncb, j = (tile % ncb) * 64 + (threadIdx.x & 63);
    if (j >= d.M + 3) return;
    uvz_cell(d, i, j, kd, kn);
}

__global__ void k_wet_count(const HdDev* __restrict__ dv)
{
    const HdDevG& d = dev_view(dv);
    int i = blockIdx.x;
    int lane = threadIdx.x;
    if (i >= d.N + 3) return;
    int c = 0;
    for (int j = lane; j < d.M + 3; j += kWave) c += (d.khouot[i * d.W + j] == 0);
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, kWave);
    /* slot 1 of the counter pair: the one-shot count */
    if (lane == 0 && c) atomicAdd((unsigned long long*)(d.wet_count + 1), (unsigned long long)c);
}

/* ================================================================ batched Thomas */

/* One system per lane.  A system outside rows [0, nrows) is skipped (its x
 * is left as it is): the host checks the tables where it holds them, and this
 * check covers tables that only the device holds. */
__global__ void k_thomas_batch(int nsys, const int* off, const int* sn, const double* A, const double* B,
                               const double* C, const double* D, double* x, double* Ap, double* Bp, int nrows)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    int o = off[s], n = sn[s];
    if (n < 0 || o < 0 || n >= nrows - o) return;
    if (n == 0) { x[o] = D[o] / B[o]; return; }
    double ap = -C[o] / B[o], bp = D[o] / B[o];
    Ap[o] = ap; Bp[o] = bp;
    for (int r = 1; r < n; ++r) {
        double e = A[o + r] * ap + B[o + r];
        double nap = -C[o + r] / e;
        double nbp = (D[o + r] - (A[o + r] * bp)) / e;
        ap = nap; bp = nbp;
        Ap[o + r] = ap; Bp[o + r] = bp;
    }
    double xv = (D[o + n] - (A[o + n] * bp)) / (B[o + n] + (A[o + n] * ap));
    x[o + n] = xv;
    for (int r = n - 1; r >= 0; --r) { xv = Bp[o + r] + (Ap[o + r] * xv); x[o + r] = xv; }
}

/* One wave per system: the speculative parallel Thomas of the run solvers
 * (hd_thomas.h) on a caller's system, rows in LDS chunks of 510, the
 * factors of earlier chunks through Ap / Bp (as k_adi_spec spills them).
 * The same bits as k_thomas_batch; systems outside rows [0, nrows) are
 * skipped the same way. */
__global__ void __launch_bounds__(64) k_thomas_spec(int nsys, const int* off, const int* sn, const double* A,
                                                    const double* B, const double* C, const double* D, double* x,
                                                    double* Ap, double* Bp, int nrows)
{
    constexpr int RW = 510;
    __shared__ double sA[RW], sB[RW], sC[RW], sD[RW];
    const int s = blockIdx.x, lane = threadIdx.x;
    if (s >= nsys) return;
    const int o = off[s], n = sn[s];
    if (n < 0 || o < 0 || n >= nrows - o) return;
    if (n == 0) {
        if (lane == 0) x[o] = D[o] / B[o];
        return;
    }
    const int nch = n / RW + 1;
    double ap = 0, bp = 0, xs = 0;
    for (int c = 0; c < nch; ++c) {
        const int rb = RW * c;
        const int re = rb + RW - 1 < n ? rb + RW - 1 : n;
        for (int i = lane; i <= re - rb; i += kWave) {
            sA[i] = A[o + rb + i]; sB[i] = B[o + rb + i]; sC[i] = C[o + rb + i]; sD[i] = D[o + rb + i];
        }
        __syncthreads();
        int g0 = 0;
        if (c == 0) {
            ap = -sC[0] / sB[0];
            bp = sD[0] / sB[0];
            g0 = 1;
        }
        const int g1 = (re < n - 1 ? re : n - 1) - rb;
        hd_spec::spec_forward<true>(sA, sB, sC, sD, g0, g1, ap, bp);
        if (c == 0) {
            __syncthreads();
            if (lane == 0) { sC[0] = -sC[0] / sB[0]; sD[0] = sD[0] / sB[0]; }
        }
        if (re == n) xs = (sD[n - rb] - (sA[n - rb] * bp)) / (sB[n - rb] + (sA[n - rb] * ap));
        __syncthreads();
        if (c + 1 < nch) {
            for (int i = lane; i <= re - rb; i += kWave) { Ap[o + rb + i] = sC[i]; Bp[o + rb + i] = sD[i]; }
            __syncthreads();
        }
    }
    if (lane == 0) x[o + n] = xs;
    double xc = xs;
    for (int c = nch - 1; c >= 0; --c) {
        const int rb = RW * c;
        const int re = rb + RW - 1 < n ? rb + RW - 1 : n;
        if (c + 1 < nch) {
            __syncthreads();
            for (int i = lane; i <= re - rb; i += kWave) { sC[i] = Ap[o + rb + i]; sD[i] = Bp[o + rb + i]; }
            __syncthreads();
        }
        const int hi = (re < n - 1 ? re : n - 1) - rb;
        xc = hd_spec::spec_backward(sC, sD, sA, hi, xc);
        __syncthreads();
        for (int i = lane; i <= hi; i += kWave) x[o + rb + i] = sA[i];
        __syncthreads();
    }
}

} // namespace

/* ================================================================ launchers */

/* Every launch covers a batch of contexts: grid.z indexes the batch, the x/y
 * extents are sized for the largest grid of the batch and each kernel
 * returns early outside its own context's extents. */
static inline dim3 grid2d(const HdBatch& b, int bx) { return dim3((b.M + 3 + bx - 1) / bx, b.N + 3, b.n); }

void hd_launch_onetime_init(const HdBatch& b, hipStream_t s)
{
    hipLaunchKernelGGL(k_onetime_init, grid2d(b, 128), dim3(128), 0, s, b.dv);
}

void hd_launch_htuongdoi(const HdBatch& b, hipStream_t s)
{
    hipLaunchKernelGGL(k_htuongdoi, dim3((b.M + 127) / 128, b.N, b.n), dim3(128), 0, s, b.dv);
}

void hd_launch_preprocess(const HdBatch& b, hipStream_t s)
{
    hipLaunchKernelGGL(k_preprocess, dim3(1, 1, b.n), dim3(64), 0, s, b.dv);
}

void hd_launch_find_limits(const HdBatch& b, hipStream_t s)
{
    hipLaunchKernelGGL(k_wet_bits, dim3((b.M + 3 + 255) / 256, b.N + 3, b.n), dim3(256), 0, s, b.dv);
    hipLaunchKernelGGL(k_reset_runs, dim3(1, 1, b.n), dim3(64), 0, s, b.dv);
    hd_launch_segments(b, 0, s);
}

/* table-scan blocks: rows 4 per block, columns 64 per block; FS edge
 * blocks: nfsr for the columns of rows N and 2, then the rows of column M */
static void scan_blocks(const HdBatch& b, int& nrowb, int& ncolb, int& nfsr, int& nfs)
{
    nrowb = (b.N + 3 + 3) / 4;
    ncolb = (b.M + 3 + 63) / 64;
    const int imax = (b.M - 1 < b.N + 2) ? b.M - 1 : b.N + 2;
    nfsr = (b.M - 1 + 255) / 256;
    nfs = nfsr + (imax + 255) / 256;
}

void hd_launch_uvz_segments(const HdBatch& b, int kd, int kn, int ntiles, int count, hipStream_t s)
{
    int nrowb, ncolb, nfsr, nfs;
    scan_blocks(b, nrowb, ncolb, nfsr, nfs);
    int ncb = (b.M + 3 + 63) / 64;
    int tiles = ntiles > 0 ? ntiles : (b.N + 3) * ncb;
    hipLaunchKernelGGL(k_uvz_segments, dim3(nrowb + ncolb + nfs + (tiles + 3) / 4, 1, b.n), dim3(256), 0, s, b.dv,
                       kd, kn, ntiles > 0 ? 1 : 0, nrowb, ncolb, nfsr, nfs, count);
}

void hd_launch_segments(const HdBatch& b, int count, hipStream_t s)
{
    int nrowb, ncolb, nfsr, nfs;
    scan_blocks(b, nrowb, ncolb, nfsr, nfs);
    hipLaunchKernelGGL(k_find_limits, dim3(nrowb + ncolb + nfs, 1, b.n), dim3(256), 0, s, b.dv, nrowb, ncolb, nfsr,
                       count);
}


void hd_launch_boundary_t(const HdBatch& b, int analytic, hipStream_t s)
{
    hipLaunchKernelGGL(k_boundary, dim3(1, 1, b.n), dim3(1024), 0, s, b.dv, analytic);
}

void hd_launch_adi_coef(const HdBatch& b, int isU, int shift, hipStream_t s)
{
    dim3 blk(64, 4);
    dim3 grd((b.M + 3 + 63) / 64, (b.N + 3 + 15) / 16, b.n);
    if (isU) hipLaunchKernelGGL(k_adi_coef<true>, grd, blk, 0, s, b.dv, shift);
    else hipLaunchKernelGGL(k_adi_coef<false>, grd, blk, 0, s, b.dv, shift);
}

void hd_launch_adi_line(const HdBatch& b, int isU, int shift, hipStream_t s)
{
    int nl = (isU ? b.M : b.N) - 1;
    if (isU) hipLaunchKernelGGL(k_adi_line<true>, dim3(nl, 1, b.n), dim3(64), 0, s, b.dv, shift);
    else hipLaunchKernelGGL(k_adi_line<false>, dim3(nl, 1, b.n), dim3(64), 0, s, b.dv, shift);
}

void hd_launch_adi_run(const HdBatch& b, int isU, int shift, int nblocks, hipStream_t s)
{
    if (isU) hipLaunchKernelGGL(k_adi_run<true>, dim3(nblocks, 1, b.n), dim3(64), 0, s, b.dv, shift);
    else hipLaunchKernelGGL(k_adi_run<false>, dim3(nblocks, 1, b.n), dim3(64), 0, s, b.dv, shift);
}

void hd_launch_adi_spec(const HdBatch& b, int isU, int shift, int nblocks, int bnd, hipStream_t s)
{
    const dim3 g(nblocks + (bnd ? 1 : 0), 1, b.n);
    if (isU) hipLaunchKernelGGL(k_adi_spec<true>, g, dim3(64), 0, s, b.dv, shift, bnd);
    else hipLaunchKernelGGL(k_adi_spec<false>, g, dim3(64), 0, s, b.dv, shift, bnd);
}

void hd_launch_adi_solve(const HdBatch& b, int isU, int shift, hipStream_t s)
{
    int nl = (isU ? b.M : b.N) - 1;
    if (isU) hipLaunchKernelGGL(k_adi_solve<true>, dim3((nl + 63) / 64, 1, b.n), dim3(64), 0, s, b.dv, shift);
    else hipLaunchKernelGGL(k_adi_solve<false>, dim3((nl + 63) / 64, 1, b.n), dim3(64), 0, s, b.dv, shift);
}

/* per-row kernels: the full grid, or ntiles active tiles (bw/64 per block) */
static inline dim3 row_grid(const HdBatch& b, int bw, int ntiles)
{
    if (ntiles > 0) return dim3((ntiles + (bw >> 6) - 1) / (bw >> 6), 1, b.n);
    return grid2d(b, bw);
}

void hd_launch_normalize_post(const HdBatch& b, int isU, int post, int horizontal, int bw, int ntiles, hipStream_t s)
{
    hipLaunchKernelGGL(k_normalize, row_grid(b, bw, ntiles), dim3(bw), 0, s, b.dv, isU, post, horizontal,
                       ntiles > 0 ? 1 : 0);
}

void hd_launch_explicit(const HdBatch& b, int isU, hipStream_t s)
{
    dim3 blk(64, 4);
    dim3 grd((b.M + 3 + 63) / 64, (b.N + 3 + 15) / 16, b.n);
    if (isU) hipLaunchKernelGGL(k_explicit<true>, grd, blk, 0, s, b.dv);
    else hipLaunchKernelGGL(k_explicit<false>, grd, blk, 0, s, b.dv);
}

void hd_launch_margins(const HdBatch& b, int isU, hipStream_t s)
{
    int nl = (isU ? b.M : b.N) - 1;
    if (isU) hipLaunchKernelGGL(k_margins<true>, dim3((nl + 63) / 64, 1, b.n), dim3(64), 0, s, b.dv);
    else hipLaunchKernelGGL(k_margins<false>, dim3((nl + 63) / 64, 1, b.n), dim3(64), 0, s, b.dv);
}

void hd_launch_explicit_margins(const HdBatch& b, int isU, int cells, hipStream_t s)
{
    int nl = (isU ? b.M : b.N) - 1;
    if (isU) hipLaunchKernelGGL(k_explicit_margins<true>, dim3(nl, 1, b.n), dim3(64), 0, s, b.dv, cells);
    else hipLaunchKernelGGL(k_explicit_margins<false>, dim3(nl, 1, b.n), dim3(64), 0, s, b.dv, cells);
}

void hd_launch_smooth_u_explicit_v(const HdBatch& b, int bw, int ntiles, int adv, hipStream_t s)
{
    hipLaunchKernelGGL(k_smooth_u_explicit_v, row_grid(b, bw, ntiles), dim3(bw), 0, s, b.dv, ntiles > 0 ? 1 : 0,
                       adv);
}

void hd_launch_smooth_v_explicit_u(const HdBatch& b, int bw, int ntiles, int adv, hipStream_t s)
{
    hipLaunchKernelGGL(k_smooth_v_explicit_u, row_grid(b, bw, ntiles), dim3(bw), 0, s, b.dv, ntiles > 0 ? 1 : 0,
                       adv);
}

void hd_launch_explicit_rows_u(const HdBatch& b, int bw, int ntiles, hipStream_t s)
{
    hipLaunchKernelGGL(k_explicit_rows_u, row_grid(b, bw, ntiles), dim3(bw), 0, s, b.dv, ntiles > 0 ? 1 : 0);
}

void hd_launch_smooth_wet(const HdBatch& b, int isU, int horizontal, int ntiles, int mode, hipStream_t s)
{
    const int nlb = (horizontal ? b.N : b.M) + 1; /* lines 1..N (1..M), one wave each */
    const int tiles = ntiles > 0 ? ntiles : (b.N + 3) * ((b.M + 3 + 63) / 64);
    hipLaunchKernelGGL(k_smooth_wet, dim3(nlb + tiles, 1, b.n), dim3(64), 0, s, b.dv, isU, horizontal,
                       ntiles > 0 ? 1 : 0, nlb, mode);
}

void hd_launch_wetting(const HdBatch& b, int horizontal, int mode, hipStream_t s)
{
    int n = (horizontal ? b.N : b.M) + 1; /* lines 1..N (1..M), one wave each */
    hipLaunchKernelGGL(k_wetting, dim3(n, 1, b.n), dim3(64), 0, s, b.dv, horizontal, mode);
}

void hd_launch_uvz_dry(const HdBatch& b, int kd, int kn, int bw, int ntiles, hipStream_t s)
{
    hipLaunchKernelGGL(k_uvz, row_grid(b, bw, ntiles), dim3(bw), 0, s, b.dv, kd, kn, ntiles > 0 ? 1 : 0);
}

void hd_launch_flow_refresh(const HdBatch& b, hipStream_t s)
{
    hipLaunchKernelGGL(k_flow_refresh, grid2d(b, 256), dim3(256), 0, s, b.dv);
}

void hd_launch_wet_count(const HdBatch& b, hipStream_t s)
{
    hipLaunchKernelGGL(k_wet_count, dim3(b.N + 3, 1, b.n), dim3(64), 0, s, b.dv);
}

int hd_thomas_batch_device(int nsys, const int* off, const int* sn, const double* A, const double* B,
                           const double* C, const double* D, double* x, double* Ap, double* Bp, int nrows,
                           hipStream_t s)
{
    if (nsys > 0)
        hipLaunchKernelGGL(k_thomas_batch, dim3((nsys + 63) / 64), dim3(64), 0, s, nsys, off, sn, A, B, C, D, x, Ap,
                           Bp, nrows);
    return 0;
}

int hd_thomas_spec_device(int nsys, const int* off, const int* sn, const double* A, const double* B,
                          const double* C, const double* D, double* x, double* Ap, double* Bp, int nrows,
                          hipStream_t s)
{
    if (nsys > 0)
        hipLaunchKernelGGL(k_thomas_spec, dim3(nsys), dim3(64), 0, s, nsys, off, sn, A, B, C, D, x, Ap, Bp, nrows);
    return 0;
}

void hd_launch_bed_extrapolate(const HdBatch& b, const double* snap, double f, hipStream_t s)
{
    long n = (long)(b.N + 3) * (b.M + 3);
    long g = (n + 255) / 256;
    hipLaunchKernelGGL(k_bed_extrapolate, dim3((unsigned)(g < 4096 ? g : 4096), 1, b.n), dim3(256), 0, s, b.dv, snap, f);
}

void hd_launch_z_change(const HdBatch& b, const double* zref, double bound, double* out, hipStream_t s)
{
    long n = (long)(b.N + 3) * (b.M + 3);
    long g = (n + 255) / 256;
    hipLaunchKernelGGL(k_z_change, dim3((unsigned)(g < 1024 ? g : 1024), 1, b.n), dim3(256), 0, s, b.dv, zref, bound, out);
}
