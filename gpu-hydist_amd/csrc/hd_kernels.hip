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

__device__ __forceinline__ unsigned long long lanemask_lt()
{
    unsigned lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

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

/* Boundary_value + FS_boundary for thread index i of one edge,
 * support_funcs.cu:544-617.  Only segment 0 of the edge line can match
 * (the `break` at support_funcs.cu:558). */
__device__ void edge_value(const HdDevG& d, int edge, int i, int t1, double t2)
{
    int N = d.N, M = d.M, W = d.W, SL = d.SL, TT = d.total_time;
    const int* bt = d.boundary_type + edge * SL;
    const int *moc, *dau, *cuoi;
    int location, loc_ext;
    bool isU;
    const double *bc, *CCs, *hi;
    double* barr;
    double hmax = d.hmax[edge];
    if (edge == 0) { isU = false; location = M; loc_ext = M + 1; moc = d.mocj; dau = d.dauj; cuoi = d.cuoij; bc = d.bc_up; CCs = d.CC_u; hi = d.hi; barr = d.vbt; }
    else if (edge == 1) { isU = false; location = 2; loc_ext = 1; moc = d.mocj; dau = d.dauj; cuoi = d.cuoij; bc = d.bc_down; CCs = d.CC_d; hi = d.hi + (N + 3); barr = d.vbd; }
    else if (edge == 2) { isU = true; location = 2; loc_ext = 1; moc = d.moci; dau = d.daui; cuoi = d.cuoii; bc = d.bc_left; CCs = d.CC_l; hi = d.hi + 2 * (N + 3); barr = d.ubt; }
    else { isU = true; location = N; loc_ext = N + 1; moc = d.moci; dau = d.daui; cuoi = d.cuoii; bc = d.bc_right; CCs = d.CC_r; hi = d.hi + 2 * (N + 3) + (M + 3); barr = d.ubp; }
    if (!(moc[location] > 0 && dau[location * SL] <= i && i <= cuoi[location * SL])) return;
    const int seg_no = 0;
    /* a column edge (0, 1) leaves its cell of row 2 / N to the row edge
     * (2 / 3) that writes the same field there: in the reference's edge
     * order the later edge wins */
    bool yield = false;
    if (!isU && (i == 2 || i == N)) {
        int re = (i == 2) ? 2 : 3;
        const int* rbt = d.boundary_type + re * SL;
        int rloc = (i == 2) ? 2 : N;
        bool covers = d.moci[rloc] > 0 && d.daui[rloc * SL] <= location && location <= d.cuoii[rloc * SL];
        yield = covers && ((rbt[0] == 2) == (bt[seg_no] == 2));
    }
    double bv = bc[seg_no * TT + t1] * (1.0 - t2) + bc[seg_no * TT + t1 + 1] * t2;
    if (bt[seg_no] == 2) {
        barr[i] = bv * hi[i];
    } else if (isU) {
        d.t_z[location * W + i] = bv;
        d.t_z[loc_ext * W + i] = bv;
    } else {
        if (!yield) d.t_z[i * W + location] = bv;
        d.t_z[i * W + loc_ext] = bv;
    }
    if (bt[seg_no] != 2 || yield) return;
    int fi, fj;
    if (isU) { fi = location; fj = i; } else { fj = location; fi = i; }
    double cv = CCs[seg_no * TT + t1] * (1.0 - t2) + CCs[seg_no * TT + t1 + 1] * t2;
    d.FS[fi * W + fj] = cv * d.htaiz_bd[fi * W + fj] * hmax;
}

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
 * The half-step time t = t_prev + dT/2 is advanced here, on the device. */
__global__ void k_boundary(const HdDev* __restrict__ dv, int analytic)
{
    const HdDevG& d = dev_view(dv);
    double* tdev = d.tdev;
    __shared__ double t_s;
    if (threadIdx.x == 0) t_s = tdev[0] + d.P.dT * 0.5;
    __syncthreads();
    double t = t_s;
    if (threadIdx.x == 0) tdev[0] = t;
    int imax = d.M > d.N ? d.M : d.N;
    if (analytic) {
        int lo = 0, hi = imax + 2;
        for (int e = 0; e < 4; ++e) {
            for (int i = lo + threadIdx.x; i <= hi; i += blockDim.x) edge_analytic(d, e, i, t);
            __syncthreads();
        }
        return;
    }
    int t1 = (int)(t / 3600);
    double t2 = (t - (3600.0 * t1)) / 3600.0;
    if (t1 < 0 || t1 + 1 >= d.total_time || t1 + 1 >= d.series_len) {
        if (threadIdx.x == 0) d.flags[1] = 1;
        return;
    }
    /* the four edges at once, 256 threads each, over the edge line's first
     * wet run (the only one the reference updates); corner cells go to the
     * later edge by the rule in edge_value */
    const int e = threadIdx.x >> 8, lt = threadIdx.x & 255;
    const int SL = d.SL;
    const int* moc = e < 2 ? d.mocj : d.moci;
    const int* dau = e < 2 ? d.dauj : d.daui;
    const int* cuoi = e < 2 ? d.cuoij : d.cuoii;
    const int location = e == 0 ? d.M : (e == 3 ? d.N : 2);
    if (moc[location] <= 0) return;
    const int a = dau[location * SL], b = cuoi[location * SL];
    for (int i = (a > 2 ? a : 2) + lt; i <= b && i <= imax; i += 256) edge_value(d, e, i, t1, t2);
}

/* ================================================================ ADI rows */

struct Pre {
    double f1, f2, f3, f5, a2, c2, d2;
};

/* _uzSolver_calculate_preindex, UVZSolver_multithread.cu:352-407, cell (i, j). */
__device__ Pre pre_uz(const HdDevG& d, int i, int j)
{
    const hd_params& P = d.P;
    int W = d.W;
    const double *u = d.u, *v = d.v, *Hm = d.H_moi, *VIS = d.VISCOIDX;
    Pre r;
    double p = 0.0, q = 0.0;
    double uc = u[i * W + j];
    double vtb = (v[i * W + j - 1] + v[i * W + j] + v[(i + 1) * W + j - 1] + v[(i + 1) * W + j]) * 0.25;
    double vis = VIS[i * W + j];
    double htu = d.Htdu[i * W + j];
    r.f1 = P.dTchia2dX * uc + vis * P.dT / P.dXbp;
    r.f2 = -(2 + d.Kx1[i * W + j] * P.dT * sqrt(uc * uc + vtb * vtb) / htu + (2 * P.dT * vis) / P.dXbp);
    r.f3 = P.dT * vis / P.dXbp - P.dTchia2dX * uc;
    if (Hm[i * W + j - 1] <= P.H_TINH) {
        if (vtb < 0) {
            p = vtb * (-3 * uc + 4 * u[i * W + j + 1] - u[i * W + j + 2]) / P.dY2;
            q = (uc - 2 * u[i * W + j + 1] + u[i * W + j + 2]) / P.dYbp;
        }
    } else {
        if (Hm[i * W + j + 1] <= P.H_TINH) {
            if ((Hm[i * W + j - 2] > P.H_TINH) && (vtb > 0)) {
                p = vtb * (3 * uc - 4 * u[i * W + j - 1] + u[i * W + j - 2]) / P.dY2;
                q = (uc - 2 * u[i * W + j - 1] + u[i * W + j - 2]) / P.dYbp;
            }
        } else {
            p = vtb * (u[i * W + j + 1] - u[i * W + j - 1]) / P.dY2;
            q = (u[i * W + j + 1] - 2 * uc + u[i * W + j - 1]) / P.dYbp;
        }
    }
    r.f5 = -2 * uc + P.dT * p - P.dT * P.CORIOLIS_FORCE * vtb - P.dT * vis * q - P.dT * (P.Windx - d.Tsxw[i * W + j]) / htu;
    r.c2 = P.dTchia2dX * htu;
    r.a2 = -P.dTchia2dX * d.Htdu[(i - 1) * W + j];
    r.d2 = d.z[i * W + j] - P.dTchia2dY * (d.Htdv[i * W + j] * v[i * W + j] - d.Htdv[i * W + j - 1] * v[i * W + j - 1]);
    return r;
}

/* _vzSolver_calculate_preindex, UVZSolver_multithread.cu:295-350, cell (i, j). */
__device__ Pre pre_vz(const HdDevG& d, int i, int j)
{
    const hd_params& P = d.P;
    int W = d.W;
    const double *u = d.u, *v = d.v, *Hm = d.H_moi, *VIS = d.VISCOIDX;
    Pre r;
    double p = 0.0, q = 0.0;
    double vc = v[i * W + j];
    double utb = (u[(i - 1) * W + j] + u[i * W + j] + u[(i - 1) * W + j + 1] + u[i * W + j + 1]) * 0.25;
    double vis = VIS[i * W + j];
    double htv = d.Htdv[i * W + j];
    r.f1 = P.dTchia2dY * vc + vis * P.dT / P.dYbp;
    r.f2 = -(2 + d.Ky1[i * W + j] * P.dT * sqrt(vc * vc + utb * utb) / htv + (2 * P.dT * vis) / P.dYbp);
    r.f3 = P.dT * vis / P.dYbp - P.dTchia2dY * vc;
    if (Hm[(i - 1) * W + j] <= P.H_TINH) {
        if (utb < 0) {
            q = utb * (-3 * vc + 4 * v[(i + 1) * W + j] - v[(i + 2) * W + j]) / P.dX2;
            p = (vc - 2 * v[(i + 1) * W + j] + v[(i + 2) * W + j]) / P.dXbp;
        }
    } else {
        if (Hm[(i + 1) * W + j] <= P.H_TINH) {
            if ((Hm[(i - 2) * W + j] > P.H_TINH) && (utb > 0)) {
                q = utb * (3 * vc - 4 * v[(i - 1) * W + j] + v[(i - 2) * W + j]) / P.dX2;
                p = (vc - 2 * v[(i - 1) * W + j] + v[(i - 2) * W + j]) / P.dXbp;
            }
        } else {
            q = utb * (v[(i + 1) * W + j] - v[(i - 1) * W + j]) / P.dX2;
            p = (v[(i + 1) * W + j] - 2 * vc + v[(i - 1) * W + j]) / P.dXbp;
        }
    }
    r.f5 = -2 * vc + P.dT * q + P.dT * P.CORIOLIS_FORCE * utb - P.dT * vis * p - P.dT * (P.Windy - d.Tsyw[i * W + j]) / htv;
    r.c2 = P.dTchia2dY * htv;
    r.a2 = -P.dTchia2dY * d.Htdv[i * W + j - 1];
    r.d2 = d.z[i * W + j] - P.dTchia2dX * (d.Htdu[i * W + j] * u[i * W + j] - d.Htdu[(i - 1) * W + j] * u[(i - 1) * W + j]);
    return r;
}

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
    for (int ri = blockIdx.x; ri < cnt; ri += gridDim.x) {
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
 * Thomas above: the same rows as k_adi_run (_calculate_preindex,
 * _calculate_abcd, _calculate_matrix_coeff, tridiag, _extract_solution,
 * UVZSolver_multithread.cu:203-723) and the same bits.  Chunks of 255 cells
 * (510 rows) are assembled by the 64 lanes, four cells each (cell t = lane +
 * 64 h; the preindex of cell t + 1 comes from the neighbouring lane by a
 * shuffle), into LDS; the forward sweep and the back substitution run over
 * the chunk on all lanes; factors of all but the last chunk go to the run's
 * slice of the Ap/Bp line buffers and come back for back substitution. */
template <bool isU>
__global__ void __launch_bounds__(64) k_adi_spec(const HdDev* __restrict__ dv, int shift)
{
    const HdDevG& d = dev_view(dv);
    const int sidx = 2 + shift, eidx = isU ? d.M - shift : d.N;
    constexpr int CH = 255; /* cells per LDS chunk (the rows of cell t need cells t and t + 1) */
    constexpr int RW = 2 * CH;
    __shared__ double sA[RW], sB[RW], sC[RW], sD[RW];  /* rows; sA holds x after the forward sweep */
    __shared__ double sAp[RW], sBp[RW];
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
    const int nbr = (lane + 1) & (kWave - 1);
    for (int ri = blockIdx.x; ri < cnt; ri += gridDim.x) {
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
                Pre pa[4] = {};
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    int k = c0 + lane + kWave * h;
                    if (k <= last) {
                        Pre a = isU ? pre_uz(d, k, line) : pre_vz(d, line, k);
                        pa[h] = a;
                        if (k == first) { sE[0] = a.a2; sE[1] = a.c2; sE[2] = a.d2; }
                        if (k == last) { sE[3] = a.a2; sE[4] = a.c2; sE[5] = a.d2; }
                    }
                }
                Pre nb[4];
#pragma unroll
                for (int h = 0; h < 4; ++h) nb[h] = shfl_pre(pa[h], nbr);
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    int t = lane + kWave * h;
                    int k = c0 + t;
                    if (t < CH && k <= last) {
                        const Pre& a = pa[h];
                        Pre b = a;
                        if (k + 1 <= last) b = (lane < kWave - 1) ? nb[h] : nb[h < 3 ? h + 1 : 3];
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
                int g0 = 0;
                if (c == 0) {
                    if (sn == 0) {
                        xs = sD[0] / sB[0];
                    } else {
                        ap = -sC[0] / sB[0];
                        bp = sD[0] / sB[0];
                        if (lane == 0) { sAp[0] = ap; sBp[0] = bp; }
                    }
                    g0 = 1;
                }
                const int g1 = (re < sn - 1 ? re : sn - 1) - rb;
                spec_forward(sA, sB, sC, sD, sAp, sBp, g0, g1, ap, bp);
                if (re == sn && sn > 0) {
                    double A = sA[sn - rb], B = sB[sn - rb], D = sD[sn - rb];
                    xs = (D - (A * bp)) / (B + (A * ap));
                }
                __syncthreads();
                if (c + 1 < nch) { /* factors of this chunk (rows rb .. re, all < sn) to the line buffers */
                    for (int i = lane; i <= re - rb; i += kWave) { gAp[rb + i] = sAp[i]; gBp[rb + i] = sBp[i]; }
                }
            }
            /* back substitution, last chunk first; x[r] replaces A[r] */
            double xc = xs;
            for (int c = nch - 1; c >= 0; --c) {
                const int rb = RW * c;
                const int re = (rb + RW - 1 < sn) ? rb + RW - 1 : sn;
                if (c + 1 < nch) {
                    __syncthreads();
                    for (int i = lane; i <= re - rb; i += kWave) { sAp[i] = gAp[rb + i]; sBp[i] = gBp[rb + i]; }
                    __syncthreads();
                } else if (lane == 0) {
                    sA[sn - rb] = xs;
                }
                const int hi = (re < sn - 1 ? re : sn - 1) - rb;
                xc = spec_backward(sAp, sBp, sA, hi, xc);
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
    double r = val[p];
    if (i >= 2 && i <= N && j >= 2 && j <= M && r != 0) {
        int count = 2;
        double nb = val[p - close] + val[p + close];
        if (d.khouot[p - far] == 0) { nb += val[p - far]; count++; }
        if (d.khouot[p + far] == 0) { nb += val[p + far]; count++; }
        r = r * d.P.heso + (1 - d.P.heso) * nb / count;
    }
    out[p] = r;
    if (!post) return;
    int kh = d.khouot[p];
    unsigned char dry = 0;
    if (i >= 1 && i <= N && j >= 1 && j <= M && kh == 0) {
        double hm = d.htaiz[p] + d.t_z[p];
        d.H_moi[p] = hm;
        bool inrange = horizontal ? (j >= 2) : (i >= 2);
        dry = inrange && (hm <= d.P.H_TINH);
    }
    dryflag[p] = dry;
    /* a position can open a neighbour only where t_z > z; t_z changes
     * during the wetting pass only at cells it opens, and those are reached
     * by the pass's own cascade */
    bool rises = d.t_z[p] > d.z[p];
    if (horizontal) {
        if (i >= 1 && i <= N && j >= 2 && j <= M && rises && (d.khouot[p - 1] == 1 || d.khouot[p + 1] == 1)) {
            int slot = atomicAdd(&cand_cnt[i], 1);
            if (slot < kCand) cand[i * kCand + slot] = j;
        }
    } else {
        if (j >= 1 && j <= M && i >= 2 && i <= N && rises && (d.khouot[p - W] == 1 || d.khouot[p + W] == 1)) {
            int slot = atomicAdd(&cand_cnt[j], 1);
            if (slot < kCand) cand[j * kCand + slot] = i;
        }
    }
}

/* ================================================================ explicit */

/* vSolver, UVZSolver_multithread.cu:727-768, at (row, col); reads the
 * smoothed t_u. */
__device__ double v_solver(const HdDevG& d, int row, int col)
{
    const hd_params& P = d.P;
    int W = d.W;
    const double *v = d.v, *u = d.u, *t_u = d.t_u, *z = d.z, *Hm = d.H_moi;
    double p, q, tmp;
    q = 0; p = 0; tmp = 0;
    double vc = v[row * W + col];
    double utb = (u[(row - 1) * W + col] + u[row * W + col] + u[(row - 1) * W + col + 1] + u[row * W + col + 1]) * 0.25;
    double t_utb = (t_u[(row - 1) * W + col] + t_u[row * W + col] + t_u[(row - 1) * W + col + 1] + t_u[row * W + col + 1]) * 0.25;
    double htv = d.Htdv[row * W + col];
    p = (v[row * W + col + 1] - v[row * W + col - 1]) / P.dY2;
    p = P.HaiChiadT + p + d.Ky1[row * W + col] * sqrt(utb * utb + vc * vc) / htv;
    if (Hm[(row - 1) * W + col] <= P.H_TINH) {
        if (utb < 0) {
            q = t_utb * (-3 * vc + 4 * v[(row + 1) * W + col] - v[(row + 2) * W + col]) / P.dX2;
            tmp = (vc - 2 * v[(row + 1) * W + col] + v[(row + 2) * W + col]) / P.dXbp;
        }
    } else {
        if (Hm[(row + 1) * W + col] <= P.H_TINH) {
            if ((Hm[(row - 2) * W + col] > P.H_TINH) && (utb > 0)) {
                q = t_utb * (3 * vc - 4 * v[(row - 1) * W + col] + v[(row - 2) * W + col]) / P.dX2;
                tmp = (vc - 2 * v[(row - 1) * W + col] + v[(row - 2) * W + col]) / P.dXbp;
            }
        } else {
            q = t_utb * (v[(row + 1) * W + col] - v[(row - 1) * W + col]) / P.dX2;
            tmp = (v[(row + 1) * W + col] - 2 * vc + v[(row - 1) * W + col]) / P.dXbp;
        }
    }
    q = P.HaiChiadT * vc - q - P.CORIOLIS_FORCE * t_utb;
    q = (q - P.g * (z[row * W + col + 1] - z[row * W + col]) / P.dY
         + d.VISCOIDX[row * W + col] * (tmp + (v[row * W + col + 1] - 2 * vc + v[row * W + col - 1]) / P.dYbp))
        + (P.Windy - d.Tsyw[row * W + col]) / htv;
    return q / p;
}

/* uSolver, UVZSolver_multithread.cu:773-815, at (row, col); reads the
 * smoothed t_v. */
__device__ double u_solver(const HdDevG& d, int row, int col)
{
    const hd_params& P = d.P;
    int W = d.W;
    const double *v = d.v, *u = d.u, *t_v = d.t_v, *z = d.z, *Hm = d.H_moi;
    double p, q, tmp;
    p = 0; q = 0; tmp = 0;
    double uc = u[row * W + col];
    double vtb = (v[row * W + col - 1] + v[row * W + col] + v[(row + 1) * W + col - 1] + v[(row + 1) * W + col]) * 0.25;
    double t_vtb = (t_v[row * W + col - 1] + t_v[row * W + col] + t_v[(row + 1) * W + col - 1] + t_v[(row + 1) * W + col]) * 0.25;
    double htu = d.Htdu[row * W + col];
    p = (u[(row + 1) * W + col] - u[(row - 1) * W + col]) / P.dX2;
    p = (P.HaiChiadT + p + d.Kx1[row * W + col] * sqrt(vtb * vtb + uc * uc) / htu);
    if (Hm[row * W + col - 1] <= P.H_TINH) {
        if (vtb < 0) {
            q = t_vtb * (-3 * uc + 4 * u[row * W + col + 1] - u[row * W + col + 2]) / P.dY2;
            tmp = (uc - 2 * u[row * W + col + 1] + u[row * W + col + 2]) / P.dYbp;
        }
    } else {
        if (Hm[row * W + col + 1] <= P.H_TINH) {
            if ((Hm[row * W + col - 2] > P.H_TINH) && (vtb > 0)) {
                q = t_vtb * (3 * uc - 4 * u[row * W + col - 1] + u[row * W + col - 2]) / P.dY2;
                tmp = (uc - 2 * u[row * W + col - 1] + u[row * W + col - 2]) / P.dYbp;
            }
        } else {
            q = t_vtb * (u[row * W + col + 1] - u[row * W + col - 1]) / P.dY2;
            tmp = (u[row * W + col + 1] - 2 * uc + u[row * W + col - 1]) / P.dYbp;
        }
    }
    q = P.HaiChiadT * uc - q + P.CORIOLIS_FORCE * t_vtb;
    q = (q - P.g * (z[(row + 1) * W + col] - z[row * W + col]) / P.dX
         + d.VISCOIDX[row * W + col] * (tmp + (u[(row + 1) * W + col] - 2 * uc + u[(row - 1) * W + col]) / P.dXbp))
        + (P.Windx - d.Tsxw[row * W + col]) / htu;
    return q / p;
}

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
__global__ void __launch_bounds__(64) k_explicit_margins(const HdDev* __restrict__ dv)
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
    for (int sg = 0; sg < nseg; ++sg) {
        const int first = dau[sg], last = cuoi[sg];
        for (int k = first + lane; k < last; k += kWave) {
            if (isU) {
                if (k >= 1 && k <= N) d.t_u[k * W + line] = u_solver(d, k, line);
            } else {
                if (k <= M && line <= N) d.t_v[line * W + k] = v_solver(d, line, k);
            }
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

/* ================================================================ wet / dry */

/* Wetting pass of Reset_states_horizontal (support_funcs.cu:92-112) /
 * _vertical (:152-171), one lane per line.  The serial scan can only act at
 * positions next to a khouot == 1 cell; those were recorded by k_normalize
 * and are visited here in increasing order, re-checking the current state, so
 * cascades along the line happen exactly as in the serial loop.  A line with
 * more candidates than fit the list is scanned in full. */
__global__ void k_wetting(const HdDev* __restrict__ dv, int horizontal)
{
    const HdDevG& d = dev_view(dv);
    int* cand_cnt = d.cand_cnt;
    int* cand = d.cand;
    unsigned char* dryflag = d.dryflag;
    /* each lane's candidate list lives in LDS (a private array would sit in
     * scratch memory, one global round trip per sort step) */
    __shared__ int s_list[64][kCand + 1];
    int* list = s_list[threadIdx.x];
    int line = blockIdx.x * blockDim.x + threadIdx.x;
    int N = d.N, M = d.M, W = d.W;
    int lo = horizontal ? 1 : 1, hi = horizontal ? N : M;
    if (line < lo || line > hi) return;
    int n = cand_cnt[line];
    if (n == 0) return;
    cand_cnt[line] = 0;
    bool full = n > kCand;
    int cnt = full ? 0 : n;
    if (!full) {
        for (int t = 0; t < n; ++t) list[t] = cand[line * kCand + t];
        for (int a = 1; a < n; ++a) { /* insertion sort */
            int v = list[a], b = a - 1;
            while (b >= 0 && list[b] > v) { list[b + 1] = list[b]; --b; }
            list[b + 1] = v;
        }
    }
    int pmax = horizontal ? M : N;
    double *tz = d.t_z, *Hm = d.H_moi, *FS = d.FS;
    const double *z = d.z, *ht = d.htaiz;
    int* kh = d.khouot;
    auto at = [&](int k) -> int { return horizontal ? line * W + k : k * W + line; };
    auto open = [&](int q, int src) {
        int pq = at(q), ps = at(src);
        tz[pq] = tz[ps];
        Hm[pq] = ht[pq] + tz[ps];
        kh[pq] = 0;
        FS[pq] = FS[ps];
        bool inrange = horizontal ? (q >= 2 && q <= M && line >= 1 && line <= N) : (q >= 2 && q <= N && line >= 1 && line <= M);
        dryflag[pq] = inrange && (Hm[pq] <= d.P.H_TINH);
    };
    /* visit positions in scan order: the recorded ones, and after a cell
     * opens its right neighbour, that neighbour (the reference scan's cascade) */
    int li = 0, pos = 1;
    bool cascade = false;
    while (true) {
        if (cascade) {
            ++pos;
        } else if (full) {
            ++pos;
        } else {
            while (li < cnt && list[li] <= pos) ++li;
            if (li >= cnt) break;
            pos = list[li];
        }
        if (pos < 2 || pos > pmax) break;
        cascade = false;
        int pp = at(pos);
        if (tz[pp] > z[pp]) {
            if (kh[at(pos - 1)] == 1) open(pos - 1, pos);
            if (kh[at(pos + 1)] == 1) { open(pos + 1, pos); cascade = true; }
        }
    }
}

/* update_uvz (support_funcs.cu:671-680) fused with the drying writes of
 * Reset_states (support_funcs.cu:114-127 / 173-185) and Htuongdoi
 * (support_funcs.cu:338-354).  Drying cell (a,b) zeroes t_u at (a-1,b),(a,b)
 * and t_v at (a,b-1),(a,b); each cell applies the flags of itself and of the
 * neighbours that write it, then copies.  Face depths read the new z (t_z)
 * inside [0,N]x[0,M] and the untouched z outside, as after update_uvz. */
__global__ void __launch_bounds__(256) k_uvz(const HdDev* __restrict__ dv, int kd, int kn, int tiled)
{
    const HdDevG& d = dev_view(dv);
    const unsigned char* dryflag = d.dryflag;
    /* the run tables are rebuilt by the next kernel on the stream */
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) { d.run_cnt[0] = 0; d.run_cnt[1] = 0; }
    int i, j;
    if (!row_cell(d, tiled, i, j)) return;
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
 * out[2] (FP64 atomics; the sums' last bits depend on the order). */
__global__ void __launch_bounds__(256) k_z_change(const HdDev* __restrict__ dv, const double* zref, double* out)
{
    const HdDevG& d = dev_view(dv);
    const int N = d.N, M = d.M, W = d.W;
    __shared__ double s1[256], s2[256];
    __shared__ unsigned long long smx[256];
    double v1 = 0, v2 = 0;
    unsigned long long mx = 0;
    const long n = (long)(N + 3) * W;
    for (long p = blockIdx.x * 256L + threadIdx.x; p < n; p += gridDim.x * 256L) {
        int i = (int)(p / W), j = (int)(p % W);
        if (i < 1 || i > N || j < 1 || j > M || d.khouot[p] != 0) continue;
        double dz = fabs(d.z[p] - zref[p]);
        unsigned long long b = __double_as_longlong(dz);
        if (!(dz <= 1e300)) b = 0x7FF0000000000000ull; /* NaN counts as infinite */
        if (b > mx) mx = b;
        v1 += d.htaiz[p] + d.z[p];
        v2 += d.htaiz[p] + zref[p];
    }
    s1[threadIdx.x] = v1;
    s2[threadIdx.x] = v2;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            s1[threadIdx.x] += s1[threadIdx.x + k];
            s2[threadIdx.x] += s2[threadIdx.x + k];
            if (smx[threadIdx.x + k] > smx[threadIdx.x]) smx[threadIdx.x] = smx[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMax((unsigned long long*)out, smx[0]);
        atomicAdd(out + 1, s1[0]);
        atomicAdd(out + 2, s2[0]);
    }
}

/* ================================================================ segments */

/* Interpolate_FS_ng, support_funcs.cu:189-208, one column j of the row
 * `location` (source row clamped to [0, N+1]; zero divisor -> quotient 0). */
__device__ void fs_ng(const HdDevG& d, int location, int offset, int sign, int j)
{
    int W = d.W;
    if (d.t_u[location * W + j] > 0) {
        double deltaX = (d.u[location * W + j] + d.t_u[location * W + j]) * 0.25 * d.P.dT;
        int i = (int)(offset + sign * (deltaX / d.P.dX));
        if (i < 0) i = 0;
        if (i > d.N + 1) i = d.N + 1;
        int i_deltaX = hd_rint_int(deltaX);
        int i_dX = hd_rint_int(d.P.dX);
        int q = i_dX != 0 ? i_deltaX / i_dX : 0;
        d.FS[location * W + j] = d.FS[i * W + j] + sign * (d.FS[(i + 1) * W + j] - d.FS[i * W + j]) * (deltaX / d.P.dX - q);
    }
}

/* Interpolate_FS_d, support_funcs.cu:210-229, one row i of column `location`
 * (source column clamped to [0, M+1]). */
__device__ void fs_d(const HdDevG& d, int location, int offset, int sign, int i)
{
    int W = d.W;
    if (d.t_v[i * W + location] > 0) {
        double deltaY = (d.v[i * W + location] + d.t_v[i * W + location]) * 0.25 * d.P.dT;
        int j = (int)(offset + sign * (deltaY / d.P.dY));
        if (j < 0) j = 0;
        if (j > d.M + 1) j = d.M + 1;
        int j_deltaY = hd_rint_int(deltaY);
        int j_dY = hd_rint_int(d.P.dY);
        int q = j_dY != 0 ? j_deltaY / j_dY : 0;
        d.FS[i * W + location] = d.FS[i * W + j + 1] + (d.FS[i * W + j + 1] - d.FS[i * W + j]) * (deltaY / d.P.dY - q);
    }
}

__global__ void k_reset_runs(const HdDev* __restrict__ dv)
{
    const HdDevG& d = dev_view(dv);
    if (threadIdx.x == 0) { d.run_cnt[0] = 0; d.run_cnt[1] = 0; }
}

/* Wet runs of one line (a row when ROW, else a column), positions 2..L, by
 * one wave: the line is read eight 64-cell chunks at a time (all loads of a
 * batch in flight together), neighbours come from lane shuffles, starts and
 * ends are compacted with ballots.  The rule is the serial scan's
 * (support_funcs.cu:252-279): a run must start at <= L-2; it ends at its last
 * wet cell, except that a run reaching L-1 or L only counts when L-2 is wet.
 * Entries k < SL go to dau / cuoi (global) and sf / sl (LDS); returns the
 * number of runs found. */
template <bool ROW>
__device__ __forceinline__ int scan_runs(const int* kh, int W, int line, int L, int SL, int lane, int* dau,
                                         int* cuoi, int* sf, int* sl)
{
    auto wetv = [&](int x) -> int {
        return (x >= 2 && x <= L) ? (kh[ROW ? line * W + x : x * W + line] == 0) : 0;
    };
    const int wLm2 = wetv(L - 2), wLm1 = wetv(L - 1);
    int ns = 0, ne = 0;
    int carry = 0; /* wet(first position of the chunk - 1) */
    constexpr int kB = 8; /* chunks of 64 cells per batch, all loads in flight together */
    for (int b0 = 2; b0 <= L; b0 += kB * kWave) {
        int w[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) w[u] = wetv(b0 + kWave * u + lane);
        const int after = wetv(b0 + kB * kWave); /* the cell after the batch */
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            if (b0 + kWave * u > L) break;
            const int x = b0 + kWave * u + lane;
            const int up = __shfl_up(w[u], 1, kWave);
            const int dn = __shfl_down(w[u], 1, kWave);
            const int nx0 = (u < kB - 1) ? __shfl(w[u < kB - 1 ? u + 1 : u], 0, kWave) : after;
            const int last_of_chunk = __shfl(w[u], kWave - 1, kWave);
            const int prevw = lane == 0 ? carry : up;
            const int nextw = lane == kWave - 1 ? nx0 : dn;
            carry = last_of_chunk;
            const bool w0 = w[u] != 0;
            const bool st = w0 && (x == 2 || !prevw) && x <= L - 2;
            const bool en = w0 && !nextw && (x <= L - 2 || (x == L - 1 && wLm2) || (x == L && wLm1 && wLm2));
            const unsigned long long sm = __ballot(st), em = __ballot(en);
            if (st) {
                int k = ns + __popcll(sm & lanemask_lt());
                if (k < SL) { dau[k] = x; sf[k] = x; }
            }
            if (en) {
                int k = ne + __popcll(em & lanemask_lt());
                if (k < SL) { cuoi[k] = x; sl[k] = x; }
            }
            ns += __popcll(sm);
            ne += __popcll(em);
        }
    }
    return ns;
}

/* Interpolate_FS_ng rows N and 2, then Interpolate_FS_d column M, by one
 * workgroup (the two reference kernels' order; the column reads the rows'
 * corner results). */
__device__ void fs_edges_block(const HdDevG& d)
{
    int N = d.N, M = d.M;
    for (int j = 1 + threadIdx.x; j < M; j += blockDim.x) {
        fs_ng(d, N, N - 1, -1, j);
    }
    for (int j = 1 + threadIdx.x; j < M; j += blockDim.x) {
        fs_ng(d, 2, 2, 1, j);
    }
    __syncthreads();
    int imax = (M - 1 < N + 2) ? M - 1 : N + 2;
    for (int i = 1 + threadIdx.x; i <= imax; i += blockDim.x) fs_d(d, M, M - 1, -1, i);
}


/* Find_Calculation_limits_Horizontal / _Vertical (support_funcs.cu:230-335):
 * the per-line tables daui / cuoii / moci (rows i = 2..N) and dauj / cuoij /
 * mocj (columns j = 2..M), plus the packed run tables the solvers read
 * (runs_v for rows, runs_u for columns; their counters were zeroed by the
 * previous kernel on the stream).  Blocks of four waves, one line per wave;
 * each block reserves its table slots with one atomic.  Blocks [0, nrowb):
 * rows 4b .. 4b+3; with count_wet these waves also add the row's
 * khouot == 0 cells (rows 0..N+2) to the wet-cell accumulator.  Blocks
 * [nrowb, nrowb + ncolb): columns 2 + 4b' ... */
__global__ void __launch_bounds__(256) k_find_limits(const HdDev* __restrict__ dv, int nrowb, int ncolb, int count_wet)
{
    const HdDevG& d = dev_view(dv);
    if (blockIdx.x == (unsigned)(nrowb + ncolb)) { /* the FS edge interpolations (independent of the mask) */
        fs_edges_block(d);
        return;
    }
    __shared__ int s_f[4][24], s_l[4][24], s_n[4], s_base[4];
    const int N = d.N, M = d.M, W = d.W, SL = d.SL;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int* kh = d.khouot;
    const bool rows = blockIdx.x < (unsigned)nrowb;
    int n = 0, line;
    if (rows) {
        const int i = 4 * blockIdx.x + wv; /* rows 0..N+2 for counting, 2..N for runs */
        line = i;
        if (count_wet && i <= N + 2) {
            int c = 0;
            for (int j = lane; j < M + 3; j += kWave) c += (kh[i * W + j] == 0);
            for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, kWave);
            if (lane == 0 && c) atomicAdd((unsigned long long*)d.wet_count, (unsigned long long)c);
        }
        if (i >= 2 && i <= N) {
            int ns = scan_runs<true>(kh, W, i, M, SL, lane, d.daui + i * SL, d.cuoii + i * SL, s_f[wv], s_l[wv]);
            n = ns < SL ? ns : SL;
            if (lane == 0) {
                d.moci[i] = n;
                if (ns > SL) d.flags[0] = 1;
            }
        }
    } else {
        const int j = 2 + 4 * (blockIdx.x - nrowb) + wv;
        line = j;
        if (j <= M) {
            int ns = scan_runs<false>(kh, W, j, N, SL, lane, d.dauj + j * SL, d.cuoij + j * SL, s_f[wv], s_l[wv]);
            n = ns < SL ? ns : SL;
            if (lane == 0) {
                d.mocj[j] = n;
                if (ns > SL) d.flags[0] = 1;
            }
        }
    }
    if (lane == 0) s_n[wv] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = s_n[0] + s_n[1] + s_n[2] + s_n[3];
        int base = tot ? atomicAdd(d.run_cnt + (rows ? 1 : 0), tot) : 0;
        for (int w = 0; w < 4; ++w) { s_base[w] = base; base += s_n[w]; }
    }
    __syncthreads();
    const int cap = rows ? (N + 3) * SL : (M + 3) * SL;
    int4* table = rows ? d.runs_v : d.runs_u;
    const int at = s_base[wv] + lane;
    if (lane < n && at < cap) table[at] = make_int4(line, s_f[wv][lane], s_l[wv][lane], 0);
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

__global__ void k_thomas_batch(int nsys, const int* off, const int* sn, const double* A, const double* B,
                               const double* C, const double* D, double* x, double* Ap, double* Bp)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    int o = off[s], n = sn[s];
    if (n < 0) return;
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
    hipLaunchKernelGGL(k_reset_runs, dim3(1, 1, b.n), dim3(64), 0, s, b.dv);
    hd_launch_segments(b, 0, s);
}

void hd_launch_segments(const HdBatch& b, int count, hipStream_t s)
{
    int nrowb = (b.N + 3 + 3) / 4, ncolb = (b.M - 1 + 3) / 4;
    /* + one block for the FS edge interpolations */
    hipLaunchKernelGGL(k_find_limits, dim3(nrowb + ncolb + 1, 1, b.n), dim3(256), 0, s, b.dv, nrowb, ncolb, count);
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

void hd_launch_adi_spec(const HdBatch& b, int isU, int shift, int nblocks, hipStream_t s)
{
    if (isU) hipLaunchKernelGGL(k_adi_spec<true>, dim3(nblocks, 1, b.n), dim3(64), 0, s, b.dv, shift);
    else hipLaunchKernelGGL(k_adi_spec<false>, dim3(nblocks, 1, b.n), dim3(64), 0, s, b.dv, shift);
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

void hd_launch_explicit_margins(const HdBatch& b, int isU, hipStream_t s)
{
    int nl = (isU ? b.M : b.N) - 1;
    if (isU) hipLaunchKernelGGL(k_explicit_margins<true>, dim3(nl, 1, b.n), dim3(64), 0, s, b.dv);
    else hipLaunchKernelGGL(k_explicit_margins<false>, dim3(nl, 1, b.n), dim3(64), 0, s, b.dv);
}

void hd_launch_wetting(const HdBatch& b, int horizontal, hipStream_t s)
{
    int n = (horizontal ? b.N : b.M) + 1;
    hipLaunchKernelGGL(k_wetting, dim3((n + 63) / 64, 1, b.n), dim3(64), 0, s, b.dv, horizontal);
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
                           const double* C, const double* D, double* x, double* Ap, double* Bp, hipStream_t s)
{
    hipLaunchKernelGGL(k_thomas_batch, dim3((nsys + 63) / 64), dim3(64), 0, s, nsys, off, sn, A, B, C, D, x, Ap, Bp);
    return 0;
}

void hd_launch_bed_extrapolate(const HdBatch& b, const double* snap, double f, hipStream_t s)
{
    long n = (long)(b.N + 3) * (b.M + 3);
    long g = (n + 255) / 256;
    hipLaunchKernelGGL(k_bed_extrapolate, dim3((unsigned)(g < 4096 ? g : 4096), 1, b.n), dim3(256), 0, s, b.dv, snap, f);
}

void hd_launch_z_change(const HdBatch& b, const double* zref, double* out, hipStream_t s)
{
    long n = (long)(b.N + 3) * (b.M + 3);
    long g = (n + 255) / 256;
    hipLaunchKernelGGL(k_z_change, dim3((unsigned)(g < 1024 ? g : 1024), 1, b.n), dim3(256), 0, s, b.dv, zref, out);
}
