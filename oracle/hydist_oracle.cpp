/*
 * hydist_oracle.cpp -- CPU FP64 restatement of the reference's ADI time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that libhydist.so is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * may load it.  It is not part of the product and the product never calls it.
 *
 * It follows the reference kernels one by one, in the order Engine.py launches
 * them (Engine.py:144-366, rep.py:126-139), on full-size arrays with the
 * reference's own scratch layout (a1..f5, AA..DD, x, Ap, Bp, ep, SN;
 * Pointers.py:47-81).  Each function cites the kernel it restates.  Where the
 * reference has a race, an out-of-bounds access or an uninitialised read, this
 * file implements one deterministic meaning, listed in the project design notes
 * as D1..D21 and repeated at the point of use.  Everything is FP64 with
 * contraction off (-ffp-contract=off); transcendental functions come from
 * hd_math.h, the same source the kernels compile.
 */
#include "../include/hydist.h"
#include "../gpu-hydist_amd/csrc/hd_math.h"

#include <cmath>
#include <cstdio>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <cstring>
#include <string>
#include <vector>

namespace {

typedef std::vector<double> vd;
typedef std::vector<int> vi;

/* Planted mistakes for the tests that show a pin catches the mistake it is
 * meant to catch (or_plant): 0 = none (always, outside those tests).  Each
 * code changes one operand, sign or index of one restated kernel. */
enum Plant {
    PL_NONE = 0,
    PL_UZ_ADV_INDEX = 1,     /* uz_preindex: central advection difference u[j+1] - u[j] */
    PL_VZ_CORIOLIS_SIGN = 2, /* vz_preindex: Coriolis term with the wrong sign */
    PL_US_ZGRAD_INDEX = 3,   /* u_solver: surface slope z[i] - z[i-1] */
    PL_VS_VISC_DROP = 4,     /* v_solver: the x-part of the viscous term dropped */
    PL_NORM_COUNT = 5,       /* normalize: neighbour count starts at 3 */
    PL_WET_OPENS_LAND = 6,   /* reset_states_h: wetting also opens khouot == 2 */
    PL_QB_AVG_INDEX = 7,     /* calculate_qb: u averaged over t_u[pos], t_u[pos-1] */
    PL_BED_DIV_SIGN = 8,     /* bed_load: x-divergence of Qbx with the wrong sign */
    PL_CHS_EROSION = 9,      /* source_chs: erosion excess divided by tob */
    PL_NCHS_GAMAC = 10,      /* source_nchs: gamac's linear coefficient dropped */
    PL_HESOK_DEPTH = 11,     /* heso_k: Kx from the node depths of the other face */
    PL_VTH_AVG = 12,         /* find_vth: u averaged over t_u[pos], t_u[pos-1] */
    PL_MARGIN_EXTRAP = 13,   /* margins_u: open-edge extrapolation 2 t[f] - t[f+2] */
    PL_UZ_FRICTION = 14,     /* uz_preindex: friction with |V|^2 instead of |V| */
    PL_UZ_D2_INDEX = 15,     /* uz_preindex: y-flux difference shifted by one column */
    PL_ABCD_C1_SIGN = 16,    /* calculate_abcd: c1's continuity term with the wrong sign */
    PL_WET_NO_CASCADE = 17,  /* reset_states_h: an opened cell does not pass the wetting on */
    PL_DRY_FACE = 18,        /* reset_states_h: drying zeroes t_u[i+1] instead of t_u[i-1] */
    PL_F4_FACTOR = 19,       /* uz_solve: surface-slope coefficient g dT / (2 dX) instead of g dT / dX */
};
int g_plant = PL_NONE;
inline bool planted(int k) { return g_plant == k; }

struct Oracle {
    hd_params P;
    int N, M, SL, W, series_len, total_time;
    double hmax_u = 0, hmax_d = 0, hmax_l = 0, hmax_r = 0;
    /* Argument_Pointers fields, constant.cuh:6-20 */
    vd h, hsnham, VISCOIDX, u, v, z, t_u, t_v, t_z, Htdu, Htdv, H_moi, htaiz, htaiz_bd,
        Kx1, Ky1, Tsxw, Tsyw, FS, tFS, VTH, Kx, Ky, Fw, Qbx, Qby, dH;
    vi khouot;
    vd ubt, ubp, vbt, vbd, hi, bc_up, bc_down, bc_left, bc_right, CC_u, CC_d, CC_l, CC_r;
    vi bienQ, boundary_type, daui, cuoii, dauj, cuoij, moci, mocj;
    /* Array_Pointers scratch, constant.cuh:22-29 */
    vd a1, b1, c1, d1, a2, c2, d2, f1, f2, f3, f5, AA, BB, CC, DD, x, Ap, Bp, ep, tmp;
    vi SN;
    int err = 0;
    std::string msg;
    bool initialised = false;

    int idx(int i, int j) const { return i * W + j; }
};

struct FieldRef {
    const char* name;
    vd Oracle::*d;
    vi Oracle::*iv;
};

const FieldRef kFields[] = {
    {"h", &Oracle::h, nullptr}, {"hsnham", &Oracle::hsnham, nullptr},
    {"VISCOIDX", &Oracle::VISCOIDX, nullptr}, {"VISCOINDX", &Oracle::VISCOIDX, nullptr},
    {"u", &Oracle::u, nullptr}, {"v", &Oracle::v, nullptr}, {"z", &Oracle::z, nullptr},
    {"t_u", &Oracle::t_u, nullptr}, {"t_v", &Oracle::t_v, nullptr}, {"t_z", &Oracle::t_z, nullptr},
    {"Htdu", &Oracle::Htdu, nullptr}, {"Htdv", &Oracle::Htdv, nullptr},
    {"H_moi", &Oracle::H_moi, nullptr}, {"htaiz", &Oracle::htaiz, nullptr},
    {"htaiz_bd", &Oracle::htaiz_bd, nullptr}, {"Kx1", &Oracle::Kx1, nullptr},
    {"Ky1", &Oracle::Ky1, nullptr}, {"Tsxw", &Oracle::Tsxw, nullptr}, {"Tsyw", &Oracle::Tsyw, nullptr},
    {"FS", &Oracle::FS, nullptr}, {"tFS", &Oracle::tFS, nullptr}, {"VTH", &Oracle::VTH, nullptr},
    {"Kx", &Oracle::Kx, nullptr}, {"Ky", &Oracle::Ky, nullptr}, {"Fw", &Oracle::Fw, nullptr},
    {"Qbx", &Oracle::Qbx, nullptr}, {"Qby", &Oracle::Qby, nullptr}, {"dH", &Oracle::dH, nullptr},
    {"khouot", nullptr, &Oracle::khouot},
    {"ubt", &Oracle::ubt, nullptr}, {"ubp", &Oracle::ubp, nullptr},
    {"vbt", &Oracle::vbt, nullptr}, {"vbd", &Oracle::vbd, nullptr}, {"hi", &Oracle::hi, nullptr},
    {"bc_up", &Oracle::bc_up, nullptr}, {"bc_down", &Oracle::bc_down, nullptr},
    {"bc_left", &Oracle::bc_left, nullptr}, {"bc_right", &Oracle::bc_right, nullptr},
    {"CC_u", &Oracle::CC_u, nullptr}, {"CC_d", &Oracle::CC_d, nullptr},
    {"CC_l", &Oracle::CC_l, nullptr}, {"CC_r", &Oracle::CC_r, nullptr},
    {"bienQ", nullptr, &Oracle::bienQ}, {"boundary_type", nullptr, &Oracle::boundary_type},
    {"daui", nullptr, &Oracle::daui}, {"cuoii", nullptr, &Oracle::cuoii},
    {"dauj", nullptr, &Oracle::dauj}, {"cuoij", nullptr, &Oracle::cuoij},
    {"moci", nullptr, &Oracle::moci}, {"mocj", nullptr, &Oracle::mocj},
    /* scratch rows of the last tridiagonal assembly (reference layout) */
    {"AA", &Oracle::AA, nullptr}, {"BB", &Oracle::BB, nullptr}, {"CC", &Oracle::CC, nullptr},
    {"DD", &Oracle::DD, nullptr}, {"x", &Oracle::x, nullptr}, {"SN", nullptr, &Oracle::SN},
    /* per-line assembly scratch (UZ: line j at j*(N+2); VZ: line i at i*(M+2)) */
    {"a1", &Oracle::a1, nullptr}, {"b1", &Oracle::b1, nullptr}, {"c1", &Oracle::c1, nullptr},
    {"d1", &Oracle::d1, nullptr}, {"a2", &Oracle::a2, nullptr}, {"c2", &Oracle::c2, nullptr},
    {"d2", &Oracle::d2, nullptr}, {"f1", &Oracle::f1, nullptr}, {"f2", &Oracle::f2, nullptr},
    {"f3", &Oracle::f3, nullptr}, {"f5", &Oracle::f5, nullptr},
};

/* ------------------------------------------------------------------ params */

/* Windx/Windy/Ufr/Dxr/wss, constant.cuh:126-133, evaluated left to right. */
void derive_params(hd_params& p)
{
    double aw = std::fabs(p.Wind);
    p.Windx = 0.0013 * (0.00075 + 0.000067 * aw) * aw * p.Wind * hd_cos(p.huonggio * (p.PI / 180));
    p.Windy = 0.0013 * (0.00075 + 0.000067 * aw) * aw * p.Wind * hd_sin(p.huonggio * (p.PI / 180));
    p.Ufr = 0.25 * hd_pow((p.Sx - 1) * p.g, 8.0 / 15.0) * hd_pow(p.dm, 9.0 / 15.0) * hd_pow(p.muy, -1.0 / 15.0);
    p.Dxr = p.dm * hd_pow(p.g * (p.Sx - 1) / p.muy * p.muy, (1.0 / 3.0));
    p.wss = 10 * p.muy / p.dm * (std::sqrt(1 + 0.01 * (p.Sx - 1) * 9.81 * hd_pow(p.dm, 3) / hd_pow(p.muy, 2)) - 1);
}

/* ------------------------------------------------------------------ init */

/* Onetime_init, support_funcs.cu:20-58.  D6: the write to row N+3 is dropped. */
void onetime_init(Oracle& o)
{
    const hd_params& P = o.P;
    int N = o.N, M = o.M, W = o.W;
    for (int i0 = 0; i0 < N + 3; ++i0) {
        for (int j0 = 0; j0 < M + 3; ++j0) {
            o.khouot[i0 * W] = 2;
            o.khouot[j0] = 2;
            int i = i0 + 1, j = j0 + 1;
            long long p = (long long)i * W + j;
            if (p < (long long)(N + 3) * W) o.khouot[p] = 2;
            if (i > N + 1 || j > M + 1) continue;
            double s4 = (o.h[o.idx(i - 1, j - 1)] + o.h[o.idx(i - 1, j)] + o.h[o.idx(i, j - 1)] + o.h[o.idx(i, j)]) * 0.25;
            if (s4 > P.NANGDAY) {
                o.khouot[o.idx(i, j)] = 0;
                o.H_moi[o.idx(i, j)] = 0;
            }
            if (i > N || j > M) continue;
            o.htaiz[o.idx(i, j)] = (o.h[o.idx(i - 1, j - 1)] + o.h[o.idx(i - 1, j)] + o.h[o.idx(i, j - 1)] + o.h[o.idx(i, j)]) * 0.25;
            o.htaiz_bd[o.idx(i, j)] = o.htaiz[o.idx(i, j)];
            if (o.h[o.idx(i, j - 1)] + o.h[o.idx(i, j)] != 0)
                o.Kx1[o.idx(i, j)] = P.g * hd_pow((o.h[o.idx(i, j - 1)] + o.h[o.idx(i, j)]) * 0.5, -2 * P.mu_mn)
                                   * hd_pow((o.hsnham[o.idx(i, j - 1)] + o.hsnham[o.idx(i, j)]) * 0.5, 2);
            if (o.h[o.idx(i - 1, j)] + o.h[o.idx(i, j)] != 0)
                o.Ky1[o.idx(i, j)] = P.g * hd_pow((o.h[o.idx(i - 1, j)] + o.h[o.idx(i, j)]) * 0.5, -2 * P.mu_mn)
                                   * hd_pow((o.hsnham[o.idx(i - 1, j)] + o.hsnham[o.idx(i, j)]) * 0.5, 2);
        }
    }
}

/* ------------------------------------------------------------------ segments */

/* Interpolate_FS_ng, support_funcs.cu:189-208.  D15: the integer division
 * i_deltaX / i_dX is kept; the source row is clamped to [0, N+1] and a zero
 * divisor gives quotient 0. */
void interpolate_fs_ng(Oracle& o, int location, int offset, int sign)
{
    const hd_params& P = o.P;
    int W = o.W;
    for (int j = 1; j < o.M; ++j) {
        if (o.t_u[location * W + j] > 0) {
            double deltaX = (o.u[location * W + j] + o.t_u[location * W + j]) * 0.25 * P.dT;
            int i = (int)(offset + sign * (deltaX / P.dX));
            if (i < 0) i = 0;
            if (i > o.N + 1) i = o.N + 1;
            int i_deltaX = hd_rint_int(deltaX);
            int i_dX = hd_rint_int(P.dX);
            int q = i_dX != 0 ? i_deltaX / i_dX : 0;
            o.FS[location * W + j] = o.FS[i * W + j] + sign * (o.FS[(i + 1) * W + j] - o.FS[i * W + j]) * (deltaX / P.dX - q);
        }
    }
}

/* Interpolate_FS_d, support_funcs.cu:210-229.  D15: the loop bound i < M is
 * kept but rows beyond N+2 (outside the array) are skipped; the source column
 * is clamped to [0, M+1]. */
void interpolate_fs_d(Oracle& o, int location, int offset, int sign)
{
    const hd_params& P = o.P;
    int W = o.W;
    for (int i = 1; i < o.M && i <= o.N + 2; ++i) {
        if (o.t_v[i * W + location] > 0) {
            double deltaY = (o.v[i * W + location] + o.t_v[i * W + location]) * 0.25 * P.dT;
            int j = (int)(offset + sign * (deltaY / P.dY));
            if (j < 0) j = 0;
            if (j > o.M + 1) j = o.M + 1;
            int j_deltaY = hd_rint_int(deltaY);
            int j_dY = hd_rint_int(P.dY);
            int q = j_dY != 0 ? j_deltaY / j_dY : 0;
            o.FS[i * W + location] = o.FS[i * W + j + 1] + (o.FS[i * W + j + 1] - o.FS[i * W + j]) * (deltaY / P.dY - q);
        }
    }
}

/* One line scan of Find_Calculation_limits_*, support_funcs.cu:252-279 /
 * 303-330.  kh(k) returns the mask along the line; `lim` is M (rows) or N
 * (columns).  D9: more than SL runs sets the overflow error instead of
 * writing past the table. */
int scan_line_q(int SL, int line, int lim, const int* kh, long long stride, int* dau, int* cuoi, int* moc)
{
    int nseg = 0, start = 2, end = 0;
    auto K = [&](int k) { return kh[k * stride]; };
    while (start < lim) {
        if (K(start) != 0) {
            while (K(start) && (start < lim)) start++;
        }
        if (start + 1 == lim) start = lim;
        if (K(start) == 0 && start + 1 < lim) {
            if (nseg >= SL) { moc[line] = SL; return -1; }
            dau[line * SL + nseg] = start;
            end = start;
            while ((K(end) == 0) && (end < lim)) end++;
            if ((K(end) != 0) && (end <= lim)) {
                cuoi[line * SL + nseg] = end - 1;
                start = end;
                nseg++;
            } else {
                cuoi[line * SL + nseg] = lim;
                start = lim;
                nseg++;
            }
        }
    }
    moc[line] = nseg;
    return nseg;
}

/* Find_Calculation_limits_Horizontal (support_funcs.cu:230-280) followed by
 * Find_Calculation_limits_Vertical (support_funcs.cu:283-335): launched back
 * to back on one stream, so the row-N / row-2 FS interpolations finish before
 * the column-M one starts. */
void find_limits(Oracle& o)
{
    int N = o.N, M = o.M, W = o.W;
    interpolate_fs_ng(o, N, N - 1, -1);
    interpolate_fs_ng(o, 2, 2, 1);
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int i = 2; i <= N; ++i)
        bad |= scan_line_q(o.SL, i, M, &o.khouot[i * W], 1, o.daui.data(), o.cuoii.data(), o.moci.data()) < 0;
    interpolate_fs_d(o, M, M - 1, -1);
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int j = 2; j <= M; ++j)
        bad |= scan_line_q(o.SL, j, N, &o.khouot[j], W, o.dauj.data(), o.cuoij.data(), o.mocj.data()) < 0;
    if (bad) { o.err = HD_ESEGOVERFLOW; o.msg = "segment_limit exceeded"; }
}

/* Htuongdoi, support_funcs.cu:338-354. */
void htuongdoi(Oracle& o)
{
    int W = o.W;
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= o.N; ++i)
        for (int j = 1; j <= o.M; ++j) {
            o.Htdu[i * W + j] = (o.h[i * W + j - 1] + o.h[i * W + j] + o.z[(i + 1) * W + j] + o.z[i * W + j]) * 0.5;
            o.Htdv[i * W + j] = (o.h[(i - 1) * W + j] + o.h[i * W + j] + o.z[i * W + j + 1] + o.z[i * W + j]) * 0.5;
        }
}

/* preprocess_data, support_funcs.cu:433-542 (single-thread serial sums).
 * D5: hi is FP64 of the exact size 2(M+N+6). */
void preprocess(Oracle& o)
{
    const hd_params& P = o.P;
    int N = o.N, M = o.M, W = o.W, SL = o.SL;
    double* hi = o.hi.data();
    const double* h = o.h.data();
    for (int k = 0; k < o.mocj[M]; ++k) {
        double sum = 0;
        for (int i = o.dauj[M * SL + k]; i <= o.cuoij[M * SL + k]; ++i) {
            hi[i] = (h[i * W + M] + h[(i - 1) * W + M]) / 2.0 - P.NANGDAY;
            sum += hd_pow(hi[i], 5.0 / 3.0);
        }
        for (int i = o.dauj[M * SL + k]; i <= o.cuoij[M * SL + k]; ++i) hi[i] = hd_pow(hi[i], 2.0 / 3.0) / (sum * P.dX);
    }
    hi += N + 3;
    for (int k = 0; k < o.mocj[2]; ++k) {
        double sum = 0;
        for (int i = o.dauj[2 * SL + k]; i <= o.cuoij[2 * SL + k]; ++i) {
            hi[i] = (h[i * W + 1] + h[(i - 1) * W + 1]) / 2.0 - P.NANGDAY;
            sum += hd_pow(hi[i], 5.0 / 3.0);
        }
        for (int i = o.dauj[2 * SL + k]; i <= o.cuoij[2 * SL + k]; ++i) hi[i] = hd_pow(hi[i], 2.0 / 3.0) / (sum * P.dX);
    }
    hi += N + 3;
    for (int k = 0; k < o.moci[2]; ++k) {
        double sum = 0;
        for (int i = o.daui[2 * SL + k]; i <= o.cuoii[2 * SL + k]; ++i) {
            hi[i] = (h[1 * W + i] + h[1 * W + i - 1]) / 2.0 - P.NANGDAY;
            sum += hd_pow(hi[i], 5.0 / 3.0);
        }
        for (int i = o.daui[2 * SL + k]; i <= o.cuoii[2 * SL + k]; ++i) hi[i] = hd_pow(hi[i], 2.0 / 3.0) / (sum * P.dX);
    }
    hi += M + 3;
    for (int k = 0; k < o.moci[N]; ++k) {
        double sum = 0;
        for (int i = o.daui[N * SL + k]; i <= o.cuoii[N * SL + k]; ++i) {
            hi[i] = (h[N * W + i] + h[N * W + i - 1]) / 2.0 - P.NANGDAY;
            sum += hd_pow(hi[i], 5.0 / 3.0);
        }
        for (int i = o.daui[N * SL + k]; i <= o.cuoii[N * SL + k]; ++i) hi[i] = hd_pow(hi[i], 2.0 / 3.0) / (sum * P.dX);
    }
    const double* hb = o.htaiz_bd.data();
    o.hmax_u = hb[2 * W + M];
    o.hmax_d = hb[2 * W + 2];
    for (int i = 3; i < N; ++i) {
        if (hb[i * W + M] > o.hmax_u) o.hmax_u = hb[i * W + M];
        if (hb[i * W + 2] > o.hmax_d) o.hmax_d = hb[i * W + 2];
    }
    o.hmax_l = hb[2 * W + 2];
    o.hmax_r = hb[N * W + 2];
    for (int i = 3; i < M; ++i) {
        if (hb[2 * W + i] > o.hmax_l) o.hmax_l = hb[2 * W + i];
        if (hb[N * W + i] > o.hmax_r) o.hmax_r = hb[N * W + i];
    }
    if (o.hmax_u != 0) o.hmax_u = 1.0 / (P.ros * o.hmax_u);
    if (o.hmax_d != 0) o.hmax_d = 1.0 / (P.ros * o.hmax_d);
    if (o.hmax_l != 0) o.hmax_l = 1.0 / (P.ros * o.hmax_l);
    if (o.hmax_r != 0) o.hmax_r = 1.0 / (P.ros * o.hmax_r);
}

/* ------------------------------------------------------------------ boundaries */

/* Boundary_value, support_funcs.cu:544-583, for every thread index i.
 * D11: the `break` sits outside the if, so only segment 0 of the edge line can
 * match. */
void boundary_value(Oracle& o, bool isU, double t, int location, int loc_ext, const int* btype, const double* hi,
                    double* barr, const vd& bc, const vi& moc, const vi& dau, const vi& cuoi)
{
    int W = o.W, SL = o.SL;
    int t1 = (int)(t / 3600);
    double t2 = (t - (3600.0 * t1)) / 3600.0;
    int imax = o.M > o.N ? o.M : o.N;
    for (int i = 2; i <= imax; ++i) {
        int seg_no = -1;
        if (moc[location] > 0 && dau[location * SL] <= i && i <= cuoi[location * SL]) seg_no = 0;
        if (seg_no == -1) continue;
        double bv = bc[seg_no * o.total_time + t1] * (1.0 - t2) + bc[seg_no * o.total_time + t1 + 1] * t2;
        if (btype[seg_no] == 2) {
            barr[i] = bv * hi[i];
        } else if (isU) {
            o.t_z[location * W + i] = bv;
            o.t_z[loc_ext * W + i] = bv;
        } else {
            o.t_z[i * W + location] = bv;
            o.t_z[i * W + loc_ext] = bv;
        }
    }
}

/* FS_boundary, support_funcs.cu:585-617 (same D11 segment rule). */
void fs_boundary(Oracle& o, bool isU, double t, int location, double hmax, const int* btype, const vd& CCs,
                 const vi& moc, const vi& dau, const vi& cuoi)
{
    int W = o.W, SL = o.SL;
    int t1 = (int)(t / 3600);
    double t2 = (t - (3600.0 * t1)) / 3600.0;
    int imax = o.M > o.N ? o.M : o.N;
    for (int k = 2; k <= imax; ++k) {
        int seg_no = -1;
        if (moc[location] > 0 && dau[location * SL] <= k && k <= cuoi[location * SL]) seg_no = 0;
        if (seg_no == -1 || btype[seg_no] != 2) continue;
        int i, j;
        if (isU) { i = location; j = k; } else { j = location; i = k; }
        double bv = CCs[seg_no * o.total_time + t1] * (1.0 - t2) + CCs[seg_no * o.total_time + t1 + 1] * t2;
        o.FS[i * W + j] = bv * o.htaiz_bd[i * W + j] * hmax;
    }
}

/* Update_Boundary_Value, support_funcs.cu:622-668: edges up, down, left,
 * right in that order (a corner cell claimed by two edges takes the later
 * edge's value). */
void update_boundary_value(Oracle& o, double t)
{
    int N = o.N, M = o.M, SL = o.SL;
    const int* bt = o.boundary_type.data();
    const double* hiu = o.hi.data();
    const double* hid = hiu + (N + 3);
    const double* hil = hid + (N + 3);
    const double* hir = hil + (M + 3);
    boundary_value(o, false, t, M, M + 1, bt, hiu, o.vbt.data(), o.bc_up, o.mocj, o.dauj, o.cuoij);
    fs_boundary(o, false, t, M, o.hmax_u, bt, o.CC_u, o.mocj, o.dauj, o.cuoij);
    boundary_value(o, false, t, 2, 1, bt + SL, hid, o.vbd.data(), o.bc_down, o.mocj, o.dauj, o.cuoij);
    fs_boundary(o, false, t, 2, o.hmax_d, bt + SL, o.CC_d, o.mocj, o.dauj, o.cuoij);
    boundary_value(o, true, t, 2, 1, bt + 2 * SL, hil, o.ubt.data(), o.bc_left, o.moci, o.daui, o.cuoii);
    fs_boundary(o, true, t, 2, o.hmax_l, bt + 2 * SL, o.CC_l, o.moci, o.daui, o.cuoii);
    boundary_value(o, true, t, N, N + 1, bt + 3 * SL, hir, o.ubp.data(), o.bc_right, o.moci, o.daui, o.cuoii);
    fs_boundary(o, true, t, N, o.hmax_r, bt + 3 * SL, o.CC_r, o.moci, o.daui, o.cuoii);
}

/* boundary_up/down/left/right, support_funcs.cu:356-429, the analytic tide of
 * the narrow-channel mode (Engine.py:30-43).  D19: bienQ is read as int32 and
 * the segment tables with stride SL; every index of the edge's first segment
 * is covered. */
void channel_boundaries(Oracle& o, double t)
{
    const hd_params& P = o.P;
    int N = o.N, M = o.M, W = o.W, SL = o.SL;
    const double PI = P.PI;
    for (int i = o.dauj[M * SL]; i <= o.cuoij[M * SL]; ++i) {
        o.t_z[i * W + M] = 0.01 * hd_cos(2 * PI / 27.750 * t) * hd_cos(2 * (PI / 100) * (100 - P.dY / 2));
        o.t_z[i * W + M + 1] = 0.01 * hd_cos(2 * PI / 27.75 * t) * hd_cos(2 * (PI / 100) * (100 + P.dY / 2));
    }
    for (int i = o.dauj[2 * SL]; i <= o.cuoij[2 * SL]; ++i) {
        if (o.bienQ[1]) o.vbd[i] = 0;
        else {
            o.t_z[i * W + 2] = 0.01 * hd_cos(2 * PI / 27.75 * t) * hd_cos(2 * (PI / 100) * P.dY / 2);
            o.t_z[i * W + 1] = 0.01 * hd_cos(2 * PI / 27.75 * t) * hd_cos(2 * (PI / 100) * (-P.dY) / 2);
        }
    }
    for (int i = o.daui[2 * SL]; i <= o.cuoii[2 * SL]; ++i) {
        if (o.bienQ[2]) o.ubt[i] = 0;
        else {
            o.t_z[2 * W + i] = 0.01 * hd_cos(2 * PI / 27.75 * t) * hd_cos(2 * (PI / 100) * P.dX / 2);
            o.t_z[1 * W + i] = 0.01 * hd_cos(2 * PI / 27.75 * t) * hd_cos(2 * (PI / 100) * (-P.dX) / 2);
        }
    }
    /* boundary_right starts at daui[2*SL] (row 2's first run), as written. */
    for (int i = o.daui[2 * SL]; i <= o.cuoii[N * SL]; ++i) {
        if (o.bienQ[3]) o.ubp[i] = 0;
        else {
            o.t_z[N * W + i] = 0.01 * hd_cos(2 * PI / 27.75 * t) * hd_cos(2 * (PI / 100) * (100 - P.dX / 2));
            o.t_z[(N + 1) * W + i] = 0.01 * hd_cos(2 * PI / 27.75 * t) * hd_cos(2 * (PI / 100) * (100 + P.dX / 2));
        }
    }
}

/* ------------------------------------------------------------------ ADI core */

/* locate_segment_u / locate_segment_v wall flags, constant.cuh:135-178. */
void wall_flags_u(const Oracle& o, int col, int first, int last, bool& br1, bool& br2)
{
    int W = o.W, N = o.N;
    const double* h = o.h.data();
    br1 = (first > 2) || ((first == 2) && ((h[1 * W + col] + h[1 * W + col - 1]) * 0.5 == o.P.NANGDAY));
    br2 = (last < N) || ((last == N) && ((h[N * W + col] + h[N * W + col - 1]) * 0.5 == o.P.NANGDAY));
}

void wall_flags_v(const Oracle& o, int row, int first, int last, bool& br1, bool& br2)
{
    int W = o.W, M = o.M;
    const double* h = o.h.data();
    br1 = (first > 2) || ((first == 2) && ((h[row * W + first - 1] + h[(row - 1) * W + first - 1]) * 0.5 == o.P.NANGDAY));
    br2 = (last < M) || ((last == M) && ((h[row * W + last] + h[(row - 1) * W + last]) * 0.5 == o.P.NANGDAY));
}

/* _uzSolver_calculate_preindex, UVZSolver_multithread.cu:352-407, cell (i, j). */
void uz_preindex(Oracle& o, int i, int j)
{
    const hd_params& P = o.P;
    int W = o.W, sw = o.N + 2;
    const double *u = o.u.data(), *v = o.v.data(), *z = o.z.data(), *Htdv = o.Htdv.data(), *Htdu = o.Htdu.data(),
                 *VIS = o.VISCOIDX.data(), *Tsxw = o.Tsxw.data(), *Hm = o.H_moi.data(), *Kx1 = o.Kx1.data();
    double* a2 = &o.a2[j * sw]; double* c2 = &o.c2[j * sw]; double* d2 = &o.d2[j * sw];
    double* f1 = &o.f1[j * sw]; double* f2 = &o.f2[j * sw]; double* f3 = &o.f3[j * sw]; double* f5 = &o.f5[j * sw];
    double p = 0.0, q = 0.0;
    double vtb = (v[i * W + j - 1] + v[i * W + j] + v[(i + 1) * W + j - 1] + v[(i + 1) * W + j]) * 0.25;
    f1[i] = P.dTchia2dX * u[i * W + j] + VIS[i * W + j] * P.dT / P.dXbp;
    double speed = std::sqrt(u[i * W + j] * u[i * W + j] + vtb * vtb);
    if (planted(PL_UZ_FRICTION)) speed = speed * speed;
    f2[i] = -(2 + Kx1[i * W + j] * P.dT * speed / Htdu[i * W + j] + (2 * P.dT * VIS[i * W + j]) / P.dXbp);
    f3[i] = P.dT * VIS[i * W + j] / P.dXbp - P.dTchia2dX * u[i * W + j];
    if (Hm[i * W + j - 1] <= P.H_TINH) {
        if (vtb < 0) {
            p = vtb * (-3 * u[i * W + j] + 4 * u[i * W + j + 1] - u[i * W + j + 2]) / P.dY2;
            q = (u[i * W + j] - 2 * u[i * W + j + 1] + u[i * W + j + 2]) / P.dYbp;
        }
    } else {
        if (Hm[i * W + j + 1] <= P.H_TINH) {
            if ((Hm[i * W + j - 2] > P.H_TINH) && (vtb > 0)) {
                p = vtb * (3 * u[i * W + j] - 4 * u[i * W + j - 1] + u[i * W + j - 2]) / P.dY2;
                q = (u[i * W + j] - 2 * u[i * W + j - 1] + u[i * W + j - 2]) / P.dYbp;
            }
        } else {
            p = vtb * (u[i * W + j + 1] - u[i * W + j - (planted(PL_UZ_ADV_INDEX) ? 0 : 1)]) / P.dY2;
            q = (u[i * W + j + 1] - 2 * u[i * W + j] + u[i * W + j - 1]) / P.dYbp;
        }
    }
    f5[i] = -2 * u[i * W + j] + P.dT * p - P.dT * P.CORIOLIS_FORCE * vtb - P.dT * VIS[i * W + j] * q
            - P.dT * (P.Windx - Tsxw[i * W + j]) / Htdu[i * W + j];
    c2[i] = P.dTchia2dX * Htdu[i * W + j];
    a2[i] = -P.dTchia2dX * Htdu[(i - 1) * W + j];
    int jd = planted(PL_UZ_D2_INDEX) ? j + 1 : j;
    d2[i] = z[i * W + j] - P.dTchia2dY * (Htdv[i * W + jd] * v[i * W + jd] - Htdv[i * W + jd - 1] * v[i * W + jd - 1]);
}

/* _vzSolver_calculate_preindex, UVZSolver_multithread.cu:295-350, cell (i, j). */
void vz_preindex(Oracle& o, int i, int j)
{
    const hd_params& P = o.P;
    int W = o.W, sw = o.M + 2;
    const double *u = o.u.data(), *v = o.v.data(), *z = o.z.data(), *Htdv = o.Htdv.data(), *Htdu = o.Htdu.data(),
                 *VIS = o.VISCOIDX.data(), *Tsyw = o.Tsyw.data(), *Hm = o.H_moi.data(), *Ky1 = o.Ky1.data();
    double* a2 = &o.a2[i * sw]; double* c2 = &o.c2[i * sw]; double* d2 = &o.d2[i * sw];
    double* f1 = &o.f1[i * sw]; double* f2 = &o.f2[i * sw]; double* f3 = &o.f3[i * sw]; double* f5 = &o.f5[i * sw];
    double p = 0.0, q = 0.0;
    double utb = (u[(i - 1) * W + j] + u[i * W + j] + u[(i - 1) * W + j + 1] + u[i * W + j + 1]) * 0.25;
    f1[j] = P.dTchia2dY * v[i * W + j] + VIS[i * W + j] * P.dT / P.dYbp;
    f2[j] = -(2 + Ky1[i * W + j] * P.dT * std::sqrt(v[i * W + j] * v[i * W + j] + utb * utb) / Htdv[i * W + j]
              + (2 * P.dT * VIS[i * W + j]) / P.dYbp);
    f3[j] = P.dT * VIS[i * W + j] / P.dYbp - P.dTchia2dY * v[i * W + j];
    if (Hm[(i - 1) * W + j] <= P.H_TINH) {
        if (utb < 0) {
            q = utb * (-3 * v[i * W + j] + 4 * v[(i + 1) * W + j] - v[(i + 2) * W + j]) / P.dX2;
            p = (v[i * W + j] - 2 * v[(i + 1) * W + j] + v[(i + 2) * W + j]) / P.dXbp;
        }
    } else {
        if (Hm[(i + 1) * W + j] <= P.H_TINH) {
            if ((Hm[(i - 2) * W + j] > P.H_TINH) && (utb > 0)) {
                q = utb * (3 * v[i * W + j] - 4 * v[(i - 1) * W + j] + v[(i - 2) * W + j]) / P.dX2;
                p = (v[i * W + j] - 2 * v[(i - 1) * W + j] + v[(i - 2) * W + j]) / P.dXbp;
            }
        } else {
            q = utb * (v[(i + 1) * W + j] - v[(i - 1) * W + j]) / P.dX2;
            p = (v[(i + 1) * W + j] - 2 * v[i * W + j] + v[(i - 1) * W + j]) / P.dXbp;
        }
    }
    f5[j] = -2 * v[i * W + j] + P.dT * q + (planted(PL_VZ_CORIOLIS_SIGN) ? -1 : 1) * P.dT * P.CORIOLIS_FORCE * utb - P.dT * VIS[i * W + j] * p
            - P.dT * (P.Windy - Tsyw[i * W + j]) / Htdv[i * W + j];
    c2[j] = P.dTchia2dY * Htdv[i * W + j];
    a2[j] = -P.dTchia2dY * Htdv[i * W + j - 1];
    d2[j] = z[i * W + j] - P.dTchia2dX * (Htdu[i * W + j] * u[i * W + j] - Htdu[(i - 1) * W + j] * u[(i - 1) * W + j]);
}

/* _calculate_abcd, UVZSolver_multithread.cu:409-466, for one segment of line
 * `line` (scratch offset line*sw).  D2: the wall overrides at first / last-1
 * win over the general formula.  The general formula is not evaluated at
 * position `last`: its value there is never part of a system. */
void calculate_abcd(Oracle& o, int line, int first, int last, double f4, int sw, bool br1, bool br2)
{
    if (first > last) return;
    double *a1 = &o.a1[line * sw], *b1 = &o.b1[line * sw], *c1 = &o.c1[line * sw], *d1 = &o.d1[line * sw];
    const double *a2 = &o.a2[line * sw], *c2 = &o.c2[line * sw], *d2 = &o.d2[line * sw];
    const double *f1 = &o.f1[line * sw], *f2 = &o.f2[line * sw], *f3 = &o.f3[line * sw], *f5 = &o.f5[line * sw];
    if (last - 1 > first) {
        for (int j = first; j < last; ++j) {
            a1[j] = f4 - f1[j] / a2[j];
            c1[j] = -f4 - (planted(PL_ABCD_C1_SIGN) ? -1 : 1) * (f3[j] / c2[j + 1]);
            b1[j] = f2[j] - (f1[j] * c2[j] / a2[j]) - (f3[j] * a2[j + 1] / c2[j + 1]);
            d1[j] = f5[j] - (f1[j] * d2[j] / a2[j]) - (f3[j] * d2[j + 1] / c2[j + 1]);
        }
        if (br1) {
            a1[first] = f4;
            b1[first] = f2[first] - (f3[first] * a2[first + 1] / c2[first + 1]);
            c1[first] = -f4 - (f3[first] / c2[first + 1]);
            d1[first] = f5[first] - (f3[first] * d2[first + 1] / c2[first + 1]);
        }
        if (br2) {
            a1[last - 1] = f4 - f1[last - 1] / a2[last - 1];
            b1[last - 1] = f2[last - 1] - (f1[last - 1] * c2[last - 1] / a2[last - 1]);
            c1[last - 1] = -f4;
            d1[last - 1] = f5[last - 1] - (f1[last - 1] * d2[last - 1] / a2[last - 1]);
        }
    } else {
        a1[first] = f4;
        b1[first] = f2[first];
        d1[first] = f5[first];
        c1[first] = -f4;
    }
}

/* bienrandau / bienlongdau, UVZSolver_multithread.cu:264-292. */
void bienrandau(int j, int first, int last, double* AA, double* BB, double* CC, double* DD, const double* a1,
                const double* b1, const double* c1, const double* d1, const double* a2, const double* c2, const double* d2)
{
    if (first > last) return;
    AA[j * 2] = a2[j]; AA[j * 2 + 1] = a1[j];
    BB[j * 2] = 1; BB[j * 2 + 1] = b1[j];
    CC[j * 2] = c2[j]; CC[j * 2 + 1] = c1[j];
    DD[j * 2] = d2[j]; DD[j * 2 + 1] = d1[j];
}

void bienlongdau(int j, int first, int last, double* AA, double* BB, double* CC, double* DD, const double* a1,
                 const double* b1, const double* c1, const double* d1, const double* a2, const double* c2, const double* d2)
{
    if (first > last) return;
    AA[j * 2] = a1[j]; AA[j * 2 + 1] = a2[j + 1];
    BB[j * 2] = b1[j]; BB[j * 2 + 1] = 1;
    CC[j * 2] = c1[j]; CC[j * 2 + 1] = c2[j + 1];
    DD[j * 2] = d1[j]; DD[j * 2 + 1] = d2[j + 1];
}

/* _calculate_matrix_coeff, UVZSolver_multithread.cu:469-593, one segment.
 * The per-thread bodies run for j = first..last; the end-row writes that the
 * reference guards with `j == last` run after all of them.  D1: offset is
 * 2*first of this segment.  D2: the end row overrides the general row. */
void calculate_matrix_coeff(Oracle& o, int line, int sw, int tw, int first, int last, int seg_no, bool br1,
                            bool br2, double ubt_or_vbd, double ubp_or_vbt, double TZ_r, double TZ_l, bool dkQ1,
                            bool dkQ2, int dkfr)
{
    if (first > last) return;
    const double *a1 = &o.a1[line * sw], *b1 = &o.b1[line * sw], *c1 = &o.c1[line * sw], *d1 = &o.d1[line * sw];
    const double *a2 = &o.a2[line * sw], *c2 = &o.c2[line * sw], *d2 = &o.d2[line * sw];
    double *AA = &o.AA[line * tw], *BB = &o.BB[line * tw], *CC = &o.CC[line * tw], *DD = &o.DD[line * tw];
    int offset = first * 2;
    int sn = 2 * (last - first);
    int end_kind = 0; /* 1: (a2,1,d2-c2*ubp), 2: (a1,b1,d1-c1*TZ_r), 3: (a2,1,d2) */
    for (int j = first; j <= last; ++j) {
        if (br1) {
            bienrandau(j, first, last, AA, BB, CC, DD, a1, b1, c1, d1, a2, c2, d2);
            DD[offset] = d2[first];
        } else {
            if (dkQ1 && (first == 2)) {
                bienrandau(j, first, last, AA, BB, CC, DD, a1, b1, c1, d1, a2, c2, d2);
                DD[offset] = d2[first] - a2[first] * ubt_or_vbd;
            } else {
                bienlongdau(j, first, last, AA, BB, CC, DD, a1, b1, c1, d1, a2, c2, d2);
                if (j == first) {
                    BB[offset] = b1[first];
                    CC[offset] = c1[first];
                    DD[offset] = d1[first] - a1[first] * TZ_l;
                }
            }
        }
    }
    /* end condition, the `j == last` branches */
    if (br1) {
        if (br2 == false) {
            if (dkQ2 && (last == dkfr)) end_kind = 1;
            else { sn--; end_kind = 2; }
        } else {
            end_kind = 3;
        }
    } else {
        bool isBienran = dkQ1 && (first == 2);
        if (br2 == false) {
            /* the reference tests dkBienQ_1 here (not dkBienQ_2), as written */
            if (dkQ1 && (last == dkfr)) {
                if (!isBienran) sn--;
                end_kind = 1;
            } else {
                sn--;
                if (!isBienran) sn -= 1;
                end_kind = 2;
            }
        } else {
            if (!isBienran) sn--;
            end_kind = 3;
        }
    }
    int r = sn + offset;
    if (end_kind == 1) { AA[r] = a2[last]; BB[r] = 1; DD[r] = d2[last] - c2[last] * ubp_or_vbt; }
    else if (end_kind == 2) { AA[r] = a1[last - 1]; BB[r] = b1[last - 1]; DD[r] = d1[last - 1] - c1[last - 1] * TZ_r; }
    else { AA[r] = a2[last]; BB[r] = 1; DD[r] = d2[last]; }
    o.SN[line * o.SL + seg_no] = sn;
}

/* tridiag, UVZSolver_multithread.cu:203-228: Thomas without pivoting on rows
 * 0..sn.  D21: sn < 0 (a one-cell run at an open start edge) has no system
 * and leaves x untouched. */
void tridiag(int sn, const double* AA, const double* BB, const double* CC, const double* DD, double* x, double* Ap,
             double* Bp, double* ep)
{
    if (sn < 0) return;
    if (sn == 0) { x[0] = DD[0] / BB[0]; return; }
    Ap[0] = -CC[0] / BB[0];
    Bp[0] = DD[0] / BB[0];
    for (int i = 1; i < sn; ++i) {
        ep[i] = AA[i] * Ap[i - 1] + BB[i];
        Ap[i] = -CC[i] / ep[i];
        Bp[i] = (DD[i] - (AA[i] * Bp[i - 1])) / ep[i];
    }
    x[sn] = (DD[sn] - (AA[sn] * Bp[sn - 1])) / (BB[sn] + (AA[sn] * Ap[sn - 1]));
    for (int i = sn - 1; i >= 0; --i) x[i] = Bp[i] + (Ap[i] * x[i + 1]);
}

/* tridiagSolver, UVZSolver_multithread.cu:231-262: every segment of one line. */
void tridiag_line(Oracle& o, int line, bool isU, int jumpstep, int tw)
{
    int nseg = isU ? o.mocj[line] : o.moci[line];
    const int* dau = isU ? o.dauj.data() : o.daui.data();
    for (int s = 0; s < nseg; ++s) {
        int first = dau[line * o.SL + s];
        int pos = line * tw + first * jumpstep + jumpstep % 2;
        tridiag(o.SN[line * o.SL + s], &o.AA[pos], &o.BB[pos], &o.CC[pos], &o.DD[pos], &o.x[pos], &o.Ap[pos],
                &o.Bp[pos], &o.ep[pos]);
    }
}

/* _uzSolver_extract_solution, UVZSolver_multithread.cu:668-723, one segment
 * of column j.  The `i == last` cell reads t_u[last-1] written by its
 * neighbour in the same launch; it runs after the other cells (intended
 * order). */
void uz_extract(Oracle& o, int j, int sn, int first, int last, bool br1, bool br2)
{
    int W = o.W, N = o.N;
    double *t_z = o.t_z.data(), *t_u = o.t_u.data();
    const double* x = &o.x[j * (2 * N + 1)];
    const double *a2 = &o.a2[j * (N + 2)], *c2 = &o.c2[j * (N + 2)], *d2 = &o.d2[j * (N + 2)];
    for (int i = first; i < last; ++i) {
        if (br1) {
            t_z[i * W + j] = x[2 * i];
            t_u[i * W + j] = x[2 * i + 1];
            t_u[(first - 1) * W + j] = 0;
        } else if (o.bienQ[2] && (first == 2)) {
            t_z[i * W + j] = x[2 * i];
            t_u[i * W + j] = x[2 * i + 1];
            t_u[(first - 1) * W + j] = o.ubt[j];
        } else {
            if (i == first) {
                t_u[first * W + j] = x[first * 2];
                t_u[(first - 1) * W + j] = (d2[first] - t_z[first * W + j] - c2[first] * t_u[first * W + j]) / a2[first];
            } else {
                t_z[i * W + j] = x[2 * i - 1];
                t_u[i * W + j] = x[2 * i];
            }
        }
    }
    int i = last;
    if (br2) {
        t_u[last * W + j] = 0;
        if (sn >= 0) t_z[last * W + j] = x[first * 2 + sn];
    } else if (o.bienQ[3] && (last == N)) {
        t_u[last * W + j] = o.ubp[j];
        if (sn >= 0) t_z[last * W + j] = x[first * 2 + sn];
    } else {
        t_u[last * W + j] = (d2[last] - a2[last] * t_u[(last - 1) * W + j] - t_z[last * W + j]) / c2[last];
    }
    (void)i;
}

/* _vzSolver_extract_solution, UVZSolver_multithread.cu:595-665, one segment
 * of row i (same ordering rule as uz_extract). */
void vz_extract(Oracle& o, int i, int sn, int first, int last, bool br1, bool br2)
{
    int W = o.W, M = o.M;
    double *t_z = o.t_z.data(), *t_v = o.t_v.data();
    const double* x = &o.x[i * (2 * M + 1)];
    const double *a2 = &o.a2[i * (M + 2)], *c2 = &o.c2[i * (M + 2)], *d2 = &o.d2[i * (M + 2)];
    for (int j = first; j < last; ++j) {
        if (br1) {
            t_z[i * W + j] = x[2 * j];
            t_v[i * W + j] = x[2 * j + 1];
            t_v[i * W + first - 1] = 0;
        } else if (o.bienQ[1] && (first == 2)) {
            t_z[i * W + j] = x[2 * j];
            t_v[i * W + j] = x[2 * j + 1];
            t_v[i * W + first - 1] = o.vbd[i];
        } else {
            if (j == first) {
                t_v[i * W + first] = x[2 * first];
                t_v[i * W + first - 1] = (d2[first] - t_z[i * W + first] - c2[first] * t_v[i * W + first]) / a2[first];
            } else {
                t_z[i * W + j] = x[2 * j - 1];
                t_v[i * W + j] = x[2 * j];
            }
        }
    }
    if (br2) {
        t_v[i * W + last] = 0;
        if (sn >= 0) t_z[i * W + last] = x[first * 2 + sn];
    } else if (o.bienQ[0] && (last == M)) {
        t_v[i * W + last] = o.vbt[i];
        if (sn >= 0) t_z[i * W + last] = x[2 * first + sn];
    } else {
        t_v[i * W + last] = (d2[last] - a2[last] * t_v[i * W + last - 1] - t_z[i * W + last]) / c2[last];
    }
}

/* UZ half-step solve, Engine.py:152-178: preindex, abcd, matrix_coeff,
 * tridiagSolver, extract over lines j in [sidx, eidx].  Every stage only
 * touches its own line's scratch and column, so running the five stages line
 * by line is the same as five grid-wide launches. */
void uz_solve(Oracle& o, int sidx, int eidx)
{
    const hd_params& P = o.P;
    int N = o.N, SL = o.SL, sw = N + 2, tw = 2 * N + 1;
    double f4 = (planted(PL_F4_FACTOR) ? 1 : 2) * P.g * P.dTchia2dX;
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = sidx; j <= eidx; ++j) {
        int nseg = o.mocj[j];
        for (int s = 0; s < nseg; ++s) {
            int first = o.dauj[j * SL + s], last = o.cuoij[j * SL + s];
            for (int i = first; i <= last; ++i) uz_preindex(o, i, j);
        }
        for (int s = 0; s < nseg; ++s) {
            int first = o.dauj[j * SL + s], last = o.cuoij[j * SL + s];
            bool br1, br2;
            wall_flags_u(o, j, first, last, br1, br2);
            calculate_abcd(o, j, first, last, f4, sw, br1, br2);
        }
        for (int s = 0; s < nseg; ++s) {
            int first = o.dauj[j * SL + s], last = o.cuoij[j * SL + s];
            bool br1, br2;
            wall_flags_u(o, j, first, last, br1, br2);
            calculate_matrix_coeff(o, j, sw, tw, first, last, s, br1, br2, o.ubt[j], o.ubp[j],
                                   o.t_z[last * o.W + j], o.t_z[first * o.W + j], o.bienQ[2] != 0, o.bienQ[3] != 0, N);
        }
        tridiag_line(o, j, true, 2, tw);
        for (int s = 0; s < nseg; ++s) {
            int first = o.dauj[j * SL + s], last = o.cuoij[j * SL + s];
            bool br1, br2;
            wall_flags_u(o, j, first, last, br1, br2);
            uz_extract(o, j, o.SN[j * SL + s], first, last, br1, br2);
        }
    }
}

/* VZ half-step solve, Engine.py:269-296, lines i in [sidx, eidx]. */
void vz_solve(Oracle& o, int sidx, int eidx)
{
    const hd_params& P = o.P;
    int M = o.M, SL = o.SL, sw = M + 2, tw = 2 * M + 1;
    double f4 = 2 * P.g * P.dTchia2dY;
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = sidx; i <= eidx; ++i) {
        int nseg = o.moci[i];
        for (int s = 0; s < nseg; ++s) {
            int first = o.daui[i * SL + s], last = o.cuoii[i * SL + s];
            for (int j = first; j <= last; ++j) vz_preindex(o, i, j);
        }
        for (int s = 0; s < nseg; ++s) {
            int first = o.daui[i * SL + s], last = o.cuoii[i * SL + s];
            bool br1, br2;
            wall_flags_v(o, i, first, last, br1, br2);
            calculate_abcd(o, i, first, last, f4, sw, br1, br2);
        }
        for (int s = 0; s < nseg; ++s) {
            int first = o.daui[i * SL + s], last = o.cuoii[i * SL + s];
            bool br1, br2;
            wall_flags_v(o, i, first, last, br1, br2);
            calculate_matrix_coeff(o, i, sw, tw, first, last, s, br1, br2, o.vbd[i], o.vbt[i],
                                   o.t_z[i * o.W + last], o.t_z[i * o.W + first], o.bienQ[1] != 0, o.bienQ[0] != 0, M);
        }
        tridiag_line(o, i, false, 2, tw);
        for (int s = 0; s < nseg; ++s) {
            int first = o.daui[i * SL + s], last = o.cuoii[i * SL + s];
            bool br1, br2;
            wall_flags_v(o, i, first, last, br1, br2);
            vz_extract(o, i, o.SN[i * SL + s], first, last, br1, br2);
        }
    }
}

/* Normalize + update_buffer, support_funcs.cu:682-729: smoothing of the
 * nonzero entries of t_u (isU) or t_v over the interior i in [2,N], j in [2,M]. */
void normalize(Oracle& o, bool isU)
{
    int N = o.N, M = o.M, W = o.W;
    int close = isU ? W : 1, far = isU ? 1 : W;
    double coeff = o.P.heso;
    vd& val = isU ? o.t_u : o.t_v;
#pragma omp parallel for schedule(static)
    for (int i = 2; i <= N; ++i)
        for (int j = 2; j <= M; ++j) {
            int p = i * W + j;
            double r = val[p];
            if (val[p] != 0) {
                int count = planted(PL_NORM_COUNT) ? 3 : 2;
                double nb = val[p - close] + val[p + close];
                if (o.khouot[p - far] == 0) { nb += val[p - far]; count++; }
                if (o.khouot[p + far] == 0) { nb += val[p + far]; count++; }
                r = r * coeff + (1 - coeff) * nb / count;
            }
            o.tmp[p] = r;
        }
#pragma omp parallel for schedule(static)
    for (int i = 2; i <= N; ++i)
        for (int j = 2; j <= M; ++j) val[i * W + j] = o.tmp[i * W + j];
}

/* vSolver, UVZSolver_multithread.cu:727-768. */
void v_solver(Oracle& o, int row, int col)
{
    const hd_params& P = o.P;
    int W = o.W;
    const double *v = o.v.data(), *u = o.u.data(), *t_u = o.t_u.data(), *z = o.z.data(), *Ky1 = o.Ky1.data(),
                 *Htdv = o.Htdv.data(), *Hm = o.H_moi.data(), *VIS = o.VISCOIDX.data(), *Tsyw = o.Tsyw.data();
    double p, q, tmp;
    q = 0; p = 0; tmp = 0;
    double utb = (u[(row - 1) * W + col] + u[row * W + col] + u[(row - 1) * W + col + 1] + u[row * W + col + 1]) * 0.25;
    double t_utb = (t_u[(row - 1) * W + col] + t_u[row * W + col] + t_u[(row - 1) * W + col + 1] + t_u[row * W + col + 1]) * 0.25;
    p = (v[row * W + col + 1] - v[row * W + col - 1]) / P.dY2;
    p = P.HaiChiadT + p + Ky1[row * W + col] * std::sqrt(utb * utb + v[row * W + col] * v[row * W + col]) / Htdv[row * W + col];
    if (Hm[(row - 1) * W + col] <= P.H_TINH) {
        if (utb < 0) {
            q = t_utb * (-3 * v[row * W + col] + 4 * v[(row + 1) * W + col] - v[(row + 2) * W + col]) / P.dX2;
            tmp = (v[row * W + col] - 2 * v[(row + 1) * W + col] + v[(row + 2) * W + col]) / P.dXbp;
        }
    } else {
        if (Hm[(row + 1) * W + col] <= P.H_TINH) {
            if ((Hm[(row - 2) * W + col] > P.H_TINH) && (utb > 0)) {
                q = t_utb * (3 * v[row * W + col] - 4 * v[(row - 1) * W + col] + v[(row - 2) * W + col]) / P.dX2;
                tmp = (v[row * W + col] - 2 * v[(row - 1) * W + col] + v[(row - 2) * W + col]) / P.dXbp;
            }
        } else {
            q = t_utb * (v[(row + 1) * W + col] - v[(row - 1) * W + col]) / P.dX2;
            tmp = (v[(row + 1) * W + col] - 2 * v[row * W + col] + v[(row - 1) * W + col]) / P.dXbp;
        }
    }
    q = P.HaiChiadT * v[row * W + col] - q - P.CORIOLIS_FORCE * t_utb;
    q = (q - P.g * (z[row * W + col + 1] - z[row * W + col]) / P.dY
         + VIS[row * W + col] * ((planted(PL_VS_VISC_DROP) ? 0.0 : tmp) + (v[row * W + col + 1] - 2 * v[row * W + col] + v[row * W + col - 1]) / P.dYbp))
        + (P.Windy - Tsyw[row * W + col]) / Htdv[row * W + col];
    o.t_v[row * W + col] = q / p;
}

/* uSolver, UVZSolver_multithread.cu:773-815. */
void u_solver(Oracle& o, int row, int col)
{
    const hd_params& P = o.P;
    int W = o.W;
    const double *v = o.v.data(), *u = o.u.data(), *t_v = o.t_v.data(), *z = o.z.data(), *Kx1 = o.Kx1.data(),
                 *Htdu = o.Htdu.data(), *Hm = o.H_moi.data(), *VIS = o.VISCOIDX.data(), *Tsxw = o.Tsxw.data();
    double p, q, tmp;
    p = 0; q = 0; tmp = 0;
    double vtb = (v[row * W + col - 1] + v[row * W + col] + v[(row + 1) * W + col - 1] + v[(row + 1) * W + col]) * 0.25;
    double t_vtb = (t_v[row * W + col - 1] + t_v[row * W + col] + t_v[(row + 1) * W + col - 1] + t_v[(row + 1) * W + col]) * 0.25;
    p = (u[(row + 1) * W + col] - u[(row - 1) * W + col]) / P.dX2;
    p = (P.HaiChiadT + p + Kx1[row * W + col] * std::sqrt(vtb * vtb + u[row * W + col] * u[row * W + col]) / Htdu[row * W + col]);
    if (Hm[row * W + col - 1] <= P.H_TINH) {
        if (vtb < 0) {
            q = t_vtb * (-3 * u[row * W + col] + 4 * u[row * W + col + 1] - u[row * W + col + 2]) / P.dY2;
            tmp = (u[row * W + col] - 2 * u[row * W + col + 1] + u[row * W + col + 2]) / P.dYbp;
        }
    } else {
        if (Hm[row * W + col + 1] <= P.H_TINH) {
            if ((Hm[row * W + col - 2] > P.H_TINH) && (vtb > 0)) {
                q = t_vtb * (3 * u[row * W + col] - 4 * u[row * W + col - 1] + u[row * W + col - 2]) / P.dY2;
                tmp = (u[row * W + col] - 2 * u[row * W + col - 1] + u[row * W + col - 2]) / P.dYbp;
            }
        } else {
            q = t_vtb * (u[row * W + col + 1] - u[row * W + col - 1]) / P.dY2;
            tmp = (u[row * W + col + 1] - 2 * u[row * W + col] + u[row * W + col - 1]) / P.dYbp;
        }
    }
    q = P.HaiChiadT * u[row * W + col] - q + P.CORIOLIS_FORCE * t_vtb;
    int zr = planted(PL_US_ZGRAD_INDEX) ? row - 1 : row;
    q = (q - P.g * (z[(zr + 1) * W + col] - z[zr * W + col]) / P.dX
         + VIS[row * W + col] * (tmp + (u[(row + 1) * W + col] - 2 * u[row * W + col] + u[(row - 1) * W + col]) / P.dXbp))
        + (P.Windx - Tsxw[row * W + col]) / Htdu[row * W + col];
    o.t_u[row * W + col] = q / p;
}

/* solveV, UVZSolver_multithread.cu:1055-1071: rows [sidx, eidx], cells of row
 * segments with first <= j < last. */
void solve_v(Oracle& o, int sidx, int eidx)
{
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = sidx; i <= eidx; ++i)
        for (int s = 0; s < o.moci[i]; ++s) {
            int first = o.daui[i * o.SL + s], last = o.cuoii[i * o.SL + s];
            if (first >= last) continue;
            for (int j = first; j < last; ++j) v_solver(o, i, j);
        }
}

/* solveU, UVZSolver_multithread.cu:947-968: rows 1..N of column segments,
 * cells first <= i < last. */
void solve_u(Oracle& o, int sidx, int eidx)
{
#pragma omp parallel for schedule(dynamic, 8)
    for (int j = sidx; j <= eidx; ++j)
        for (int s = 0; s < o.mocj[j]; ++s) {
            int first = o.dauj[j * o.SL + s], last = o.cuoij[j * o.SL + s];
            if (first > last) continue;
            for (int i = first; i < last; ++i)
                if (i >= 1 && i <= o.N) u_solver(o, i, j);
        }
}

const double kEpsilon = 1e-16; /* UVZSolver_multithread.cu:12 */

/* update_margin_elem_V, UVZSolver_multithread.cu:1008-1052. */
void margins_v(Oracle& o, int sidx, int eidx)
{
    int W = o.W, M = o.M;
    const double* h = o.h.data();
    double* t_v = o.t_v.data();
    for (int i = sidx; i <= eidx; ++i)
        for (int k = 0; k < o.moci[i]; ++k) {
            int first = o.daui[i * o.SL + k], last = o.cuoii[i * o.SL + k];
            bool br1 = (first > 2) || ((first == 2) && (std::fabs((h[i * W + first - 1] + h[(i - 1) * W + first - 1]) * 0.5 - o.P.NANGDAY) < kEpsilon));
            bool br2 = (last < M) || ((last == M) && (std::fabs((h[i * W + last] + h[(i - 1) * W + last]) * 0.5 - o.P.NANGDAY) < kEpsilon));
            if (br1) t_v[i * W + first - 1] = 0;
            else t_v[i * W + first - 1] = 2 * t_v[i * W + first] - t_v[i * W + first + 1];
            if (br2) t_v[i * W + last] = 0;
            else t_v[i * W + last] = 2 * t_v[i * W + last - 1] - t_v[i * W + last - 2];
        }
}

/* update_margin_elem_U, UVZSolver_multithread.cu:970-1006 (note the abs over
 * the sum only at the first edge, :989, as written). */
void margins_u(Oracle& o, int sidx, int eidx)
{
    int W = o.W, N = o.N;
    const double* h = o.h.data();
    double* t_u = o.t_u.data();
    for (int j = sidx; j <= eidx; ++j)
        for (int k = 0; k < o.mocj[j]; ++k) {
            int first = o.dauj[j * o.SL + k], last = o.cuoij[j * o.SL + k];
            bool br1 = (first > 2) || ((first == 2) && ((std::fabs(h[1 * W + j] + h[1 * W + j - 1]) * 0.5 - o.P.NANGDAY) < kEpsilon));
            bool br2 = (last < N) || ((last == N) && (std::fabs((h[N * W + j] + h[N * W + j - 1]) * 0.5 - o.P.NANGDAY) < kEpsilon));
            if (br1) t_u[(first - 1) * W + j] = 0;
            else t_u[(first - 1) * W + j] = 2 * t_u[first * W + j] - t_u[(first + (planted(PL_MARGIN_EXTRAP) ? 2 : 1)) * W + j];
            if (br2) t_u[last * W + j] = 0;
            else t_u[last * W + j] = 2 * t_u[(last - 1) * W + j] - t_u[(last - 2) * W + j];
        }
}

/* update_h_moi, support_funcs.cu:60-73. */
void update_h_moi(Oracle& o)
{
    int W = o.W;
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= o.N; ++i)
        for (int j = 1; j <= o.M; ++j)
            if (o.khouot[i * W + j] == 0) o.H_moi[i * W + j] = o.htaiz[i * W + j] + o.t_z[i * W + j];
}

/* Reset_states_horizontal, support_funcs.cu:75-129 (serial along each row). */
void reset_states_h(Oracle& o)
{
    const hd_params& P = o.P;
    int W = o.W, M = o.M;
    double *Hm = o.H_moi.data(), *htaiz = o.htaiz.data(), *t_z = o.t_z.data(), *t_u = o.t_u.data(), *t_v = o.t_v.data(),
           *FS = o.FS.data();
    const double* z = o.z.data();
    int* kh = o.khouot.data();
    std::vector<int> opened; /* PL_WET_NO_CASCADE only */
    auto opened_here = [&](int p) {
        for (int q : opened) if (q == p) return true;
        return false;
    };
    for (int i = 1; i <= o.N; ++i) {
        for (int j = 2; j <= M; ++j) {
            if (planted(PL_WET_NO_CASCADE) && opened_here(i * W + j)) continue;
            if (t_z[i * W + j] > z[i * W + j]) {
                if (kh[i * W + j - 1] == 1 || (planted(PL_WET_OPENS_LAND) && kh[i * W + j - 1] == 2)) {
                    t_z[i * W + j - 1] = t_z[i * W + j];
                    Hm[i * W + j - 1] = htaiz[i * W + j - 1] + t_z[i * W + j];
                    kh[i * W + j - 1] = 0;
                    FS[i * W + j - 1] = FS[i * W + j];
                }
                if (kh[i * W + j + 1] == 1 || (planted(PL_WET_OPENS_LAND) && kh[i * W + j + 1] == 2)) {
                    t_z[i * W + j + 1] = t_z[i * W + j];
                    Hm[i * W + j + 1] = htaiz[i * W + j + 1] + t_z[i * W + j];
                    kh[i * W + j + 1] = 0;
                    FS[i * W + j + 1] = FS[i * W + j];
                    if (planted(PL_WET_NO_CASCADE)) opened.push_back(i * W + j + 1);
                }
            }
        }
        for (int j = 2; j <= M; ++j) {
            if ((kh[i * W + j] == 0) && (Hm[i * W + j] <= P.H_TINH)) {
                t_u[(planted(PL_DRY_FACE) ? i + 1 : i - 1) * W + j] = 0;
                t_u[i * W + j] = 0;
                t_v[i * W + j - 1] = 0;
                t_v[i * W + j] = 0;
                kh[i * W + j] = 1;
                FS[i * W + j] = 0;
            }
        }
    }
}

/* Reset_states_vertical, support_funcs.cu:134-186 (serial along each column). */
void reset_states_v(Oracle& o)
{
    const hd_params& P = o.P;
    int W = o.W, N = o.N;
    double *Hm = o.H_moi.data(), *htaiz = o.htaiz.data(), *t_z = o.t_z.data(), *t_u = o.t_u.data(), *t_v = o.t_v.data(),
           *FS = o.FS.data();
    const double* z = o.z.data();
    int* kh = o.khouot.data();
    for (int j = 1; j <= o.M; ++j) {
        for (int i = 2; i <= N; ++i) {
            if (t_z[i * W + j] > z[i * W + j]) {
                if (kh[(i - 1) * W + j] == 1) {
                    t_z[(i - 1) * W + j] = t_z[i * W + j];
                    Hm[(i - 1) * W + j] = htaiz[(i - 1) * W + j] + t_z[i * W + j];
                    kh[(i - 1) * W + j] = 0;
                    FS[(i - 1) * W + j] = FS[i * W + j];
                }
                if (kh[(i + 1) * W + j] == 1) {
                    t_z[(i + 1) * W + j] = t_z[i * W + j];
                    Hm[(i + 1) * W + j] = htaiz[(i + 1) * W + j] + t_z[i * W + j];
                    kh[(i + 1) * W + j] = 0;
                    FS[(i + 1) * W + j] = FS[i * W + j];
                }
            }
        }
        for (int i = 2; i <= N; ++i) {
            if ((kh[i * W + j] == 0) && (Hm[i * W + j] <= P.H_TINH)) {
                t_u[(i - 1) * W + j] = 0;
                t_u[i * W + j] = 0;
                t_v[i * W + j - 1] = 0;
                t_v[i * W + j] = 0;
                kh[i * W + j] = 1;
                FS[i * W + j] = 0;
            }
        }
    }
}

/* update_uvz, support_funcs.cu:671-680.  D10: the channel flags are the
 * intended kernel arguments (u = t_u (1-kenhhepd), v = t_v (1-kenhhepng)). */
void update_uvz(Oracle& o, int kd, int kn)
{
    int W = o.W;
#pragma omp parallel for schedule(static)
    for (int i = 0; i <= o.N; ++i)
        for (int j = 0; j <= o.M; ++j) {
            o.z[i * W + j] = o.t_z[i * W + j];
            o.u[i * W + j] = o.t_u[i * W + j] * (1 - kd);
            o.v[i * W + j] = o.t_v[i * W + j] * (1 - kn);
        }
}

/* ------------------------------------------------------------------ sediment */

/* source_chs, Sediment_Transport.cu:12-49. */
double source_chs(const hd_params& P, double wsm, double Zf, double fs, double fw, double vth, double h_moi)
{
    double toee = P.Toe, todd = P.Tod, S = 0;
    double Pd, beta, Cb;
    double tob = P.ro * fw * (vth * vth);
    if (h_moi > P.ghtoe) toee = P.Toe * (1 + P.hstoe * ((h_moi - P.ghtoe)));
    if (tob > toee) S = P.Mbochat / P.ros * (tob - toee) / (planted(PL_CHS_EROSION) ? tob : toee);
    else if (tob < todd) {
        Pd = 1 - (tob / todd);
        beta = 1 + (Zf / (1.25 + 4.75 * hd_pow(Pd, 2.5)));
        Cb = beta * fs;
        S = -wsm * Cb * Pd;
    }
    return S;
}

/* source_nchs, Sediment_Transport.cu:52-60. */
double source_nchs(const hd_params& P, double wsm, double Zf, double Dxr, double Ufr, double Uf, double h_moi, double fs)
{
    double gamac = 0.434 + (planted(PL_NCHS_GAMAC) ? 0.0 : 5.975) * Zf + 2.888 * Zf * Zf;
    double Tx = std::fmax(0.0, (Uf * Uf - Ufr * Ufr) / (Ufr * Ufr));
    double Cac = 0.015 * P.dm * hd_pow(Tx, 1.5) * hd_pow(Dxr, (-0.3)) / (0.05 * h_moi);
    return wsm * (Cac - gamac * fs);
}

/* Find_VTH, Sediment_Transport.cu:541-568. */
void find_vth(Oracle& o)
{
    int W = o.W;
    for (int i = 2; i <= o.N; ++i)
        for (int j = 2; j <= o.M; ++j) {
            int pos = i * W + j;
            if ((o.htaiz[pos] > o.P.NANGDAY) && (o.H_moi[pos] > o.P.H_TINH)) {
                double ut = (o.t_u[pos] + o.t_u[pos - (planted(PL_VTH_AVG) ? 1 : W)]) * 0.5;
                double vt = (o.t_v[pos] + o.t_v[pos - 1]) * 0.5;
                o.VTH[pos] = std::sqrt(ut * ut + vt * vt);
            } else o.VTH[pos] = 0;
        }
}

/* hesoK, Sediment_Transport.cu:508-539.  The launch also covers row 0 and
 * column 0, where Htdu/Htdv are never written (Htuongdoi covers [1,N]x[1,M])
 * and stay zero, so those threads write nothing; the loop starts at 1. */
void heso_k(Oracle& o)
{
    const hd_params& P = o.P;
    int W = o.W;
    for (int i = 1; i <= o.N; ++i)
        for (int j = 1; j <= o.M; ++j) {
            int pos = i * W + j;
            double Cz;
            if (o.Htdu[pos] > 0) {
                Cz = 18 * hd_log(12 * o.Htdu[pos] / P.Ks);
                o.Kx[pos] = 5.93 * std::sqrt(P.g) * (o.h[pos - (planted(PL_HESOK_DEPTH) ? W : 1)] + o.h[pos]) * 0.25 * std::fabs(o.t_u[pos]) / Cz;
                o.Kx[pos] = std::fmin(100.0, std::fmax(5.0, o.Kx[pos]));
            }
            if (o.Htdv[pos] > 0) {
                Cz = 18 * hd_log(12 * o.Htdv[pos] / P.Ks);
                o.Ky[pos] = 5.93 * std::sqrt(P.g) * (o.h[pos - W] + o.h[pos]) * 0.25 * std::fabs(o.t_v[pos]) / Cz;
                o.Ky[pos] = std::fmin(100.0, std::fmax(5.0, o.Ky[pos]));
            }
        }
}

/* _FSj_calculate__mactrix_coeff, Sediment_Transport.cu:205-302, one cell of a
 * column segment.  D14: gamav is evaluated before Zf and the last row uses &&,
 * as written. */
void fsj_matrix_cell(Oracle& o, double t, double s_start, bool ketdinh, int i, int j, int first, int last, bool br1, bool br2)
{
    const hd_params& P = o.P;
    int W = o.W, tw = o.N + 3;
    double *AA = &o.AA[j * tw], *BB = &o.BB[j * tw], *CC = &o.CC[j * tw], *DD = &o.DD[j * tw];
    const double *FS = o.FS.data(), *Hm = o.H_moi.data(), *t_u = o.t_u.data(), *t_v = o.t_v.data(), *VTH = o.VTH.data(),
                 *Htdu = o.Htdu.data(), *Htdv = o.Htdv.data(), *Fw = o.Fw.data(), *Kx = o.Kx.data(), *Ky = o.Ky.data();
    int pos = i * W + j;
    double c = 18 * hd_log(12 * Hm[pos] / P.Ks);
    double Uf = std::sqrt(P.g) * std::fabs(VTH[pos]) / c;
    double wsm = P.wss * hd_pow((1 - FS[pos]), 4);
    double Zf = 0, S = 0;
    double gamav = 0.98 - 0.198 * Zf + 0.032 * Zf * Zf;
    if (FS[pos] > 0.00001) Zf = wsm / (0.4 * (Uf + 2 * wsm));
    if (ketdinh) S = source_chs(P, wsm, Zf, FS[pos], Fw[pos], VTH[pos], Hm[pos]);
    else S = source_nchs(P, wsm, Zf, P.Dxr, P.Ufr, Uf, Hm[pos], FS[pos]);
    if (t < s_start) S = 0;
    AA[i] = -gamav * 0.5 * (t_u[pos] + t_u[pos - W]) / P.dX2 - Htdu[pos - W] * Kx[pos - W] / (Hm[pos] * (P.dXbp));
    CC[i] = gamav * 0.5 * (t_u[pos] + t_u[pos - W]) / P.dX2 - Htdu[pos] * Kx[pos] / (Hm[pos] * (P.dXbp));
    BB[i] = P.HaiChiadT + (Htdu[pos] * Kx[pos] + Htdu[pos - W] * Kx[pos - W]) / (Hm[pos] * (P.dXbp));
    double p = 0, q = 0;
    if ((Hm[pos - 1] > P.H_TINH) && (Hm[pos + 2] > P.H_TINH)) p = (FS[pos + 1] - FS[pos]);
    if ((Hm[pos - 2] > P.H_TINH) && (Hm[pos + 1] > P.H_TINH)) q = (FS[pos] - FS[pos - 1]);
    p = (1 / (Hm[pos] * P.dYbp)) * (Htdv[pos] * Ky[pos] * p - Htdv[pos - 1] * Ky[pos - 1] * q);
    if ((Hm[pos - 2] > P.H_TINH) && (Hm[pos + 2] > P.H_TINH)) {
        q = (FS[pos + 1] - FS[pos - 1]) / P.dY2;
        q = (t_v[pos] + t_v[pos - 1]) * 0.5 * q * gamav;
    } else q = 0;
    DD[i] = FS[pos] / (P.dT * 0.5) - q + p + (S / Hm[pos]);
    if (i == first + 1) {
        if ((br1) || (t_u[first * W + j] == 0)) BB[i] += AA[i];
        else DD[i] -= AA[i] * FS[first * W + j];
    }
    if (i == last - 1) {
        if ((br2) && (t_u[last * W + j] == 0)) BB[i] += CC[i];
        else DD[i] -= CC[i] * FS[last * W + j];
    }
}

/* _FSi_calculate__mactrix_coeff, Sediment_Transport.cu:62-165, one cell of a
 * row segment (gamav after Zf, || at the last row, as written). */
void fsi_matrix_cell(Oracle& o, double t, double s_start, bool ketdinh, int i, int j, int first, int last, bool br1, bool br2)
{
    const hd_params& P = o.P;
    int W = o.W, tw = o.M + 3;
    double *AA = &o.AA[i * tw], *BB = &o.BB[i * tw], *CC = &o.CC[i * tw], *DD = &o.DD[i * tw];
    const double *FS = o.FS.data(), *Hm = o.H_moi.data(), *t_u = o.t_u.data(), *t_v = o.t_v.data(), *VTH = o.VTH.data(),
                 *Htdu = o.Htdu.data(), *Htdv = o.Htdv.data(), *Fw = o.Fw.data(), *Kx = o.Kx.data(), *Ky = o.Ky.data();
    int pos = i * W + j;
    double c = 18 * hd_log(12 * Hm[pos] / P.Ks);
    double Uf = std::sqrt(P.g) * std::fabs(VTH[pos]) / c;
    double wsm = P.wss * hd_pow((1 - FS[pos]), 4);
    double Zf = 0;
    if (FS[pos] > 0.00001) Zf = wsm / (0.4 * (Uf + 2 * wsm));
    double S;
    if (ketdinh) S = source_chs(P, wsm, Zf, FS[pos], Fw[pos], VTH[pos], Hm[pos]);
    else S = source_nchs(P, wsm, Zf, P.Dxr, P.Ufr, Uf, Hm[pos], FS[pos]);
    if (t < s_start) S = 0;
    double gamav = 0.98 - 0.198 * Zf + 0.032 * Zf * Zf;
    AA[j] = -gamav * 0.5 * (t_v[pos - 1] + t_v[pos]) / P.dY2 - Htdv[pos - 1] * Ky[pos - 1] / (Hm[pos] * P.dYbp);
    CC[j] = gamav * 0.5 * (t_v[pos - 1] + t_v[pos]) / P.dY2 - Htdv[pos] * Ky[pos] / (Hm[pos] * P.dYbp);
    BB[j] = P.HaiChiadT + (Htdv[pos] * Ky[pos] + Htdv[pos - 1] * Ky[pos - 1]) / (Hm[pos] * P.dYbp);
    double p = 0, q = 0;
    if ((Hm[pos + 2 * W] > P.H_TINH) && (Hm[pos - W] > P.H_TINH)) p = (FS[pos + W] - FS[pos]);
    if (Hm[pos - 2 * W] > P.H_TINH && Hm[pos + W] > P.H_TINH) q = (FS[pos] - FS[pos - W]);
    p = (1 / (Hm[pos] * P.dXbp)) * (Htdu[pos] * Kx[pos] * p - Htdu[pos - W] * Kx[pos - W] * q);
    if (Hm[pos + 2 * W] > P.H_TINH && Hm[pos - 2 * W] > P.H_TINH) {
        q = (FS[pos + W] - FS[pos - W]) / P.dX2;
        q = (t_u[pos] + t_u[pos - W]) * 0.5 * q * gamav;
    } else q = 0;
    DD[j] = FS[pos] / (P.dT * 0.5) - q + p + (S / Hm[pos]);
    if (j == first + 1) {
        if ((br1) || (t_v[i * W + first] == 0)) BB[j] = BB[j] + AA[j];
        else DD[j] = DD[j] - AA[j] * FS[i * W + first];
    }
    if (j == last - 1) {
        if ((br2) || (t_v[i * W + last] == 0)) BB[j] = BB[j] + CC[j];
        else DD[j] = DD[j] - CC[j] * FS[i * W + last];
    }
}

/* Scan_FSj + tridiagSolver(jump 1) + FSj_extract_solution + Update_FS,
 * Engine.py:218-239 (Sediment_Transport.cu:594-629).  Segments with
 * last - first < 2 have no rows: their tridiagSolver call works on a stale SN
 * and writes only x entries that no extraction reads, so they are skipped. */
void sediment_j(Oracle& o, double t, double s_start, bool ketdinh)
{
    int W = o.W, N = o.N, M = o.M, SL = o.SL, tw = N + 3;
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = 3; j <= M - 1; ++j) {
        for (int s = 0; s < o.mocj[j]; ++s) {
            int first = o.dauj[j * SL + s], last = o.cuoij[j * SL + s];
            if (last < first + 2) continue;
            bool br1, br2;
            wall_flags_u(o, j, first, last, br1, br2);
            for (int i = first + 1; i <= last - 1; ++i) fsj_matrix_cell(o, t, s_start, ketdinh, i, j, first, last, br1, br2);
            int sn = last - first - 2;
            o.SN[j * SL + s] = sn;
            int pos = j * tw + first + 1;
            tridiag(sn, &o.AA[pos], &o.BB[pos], &o.CC[pos], &o.DD[pos], &o.x[pos], &o.Ap[pos], &o.Bp[pos], &o.ep[pos]);
        }
    }
    for (int j = 3; j <= M - 1; ++j) {
        for (int s = 0; s < o.mocj[j]; ++s) {
            int first = o.dauj[j * SL + s], last = o.cuoij[j * SL + s];
            if (last < first + 2) continue;
            bool br1, br2;
            wall_flags_u(o, j, first, last, br1, br2);
            const double* x = &o.x[j * tw];
            for (int i = first + 1; i <= last - 1; ++i) {
                int pos = i * W + j;
                o.tFS[pos] = x[i];
                if (x[i] < 0) o.tFS[pos] = o.P.NDnen;
                if (i == first + 1) {
                    if ((br1) || (o.t_u[pos - W] == 0)) o.tFS[first * W + j] = o.tFS[pos];
                    else o.tFS[first * W + j] = o.FS[first * W + j];
                }
                if (i == last - 1) {
                    if ((br2) || (o.t_u[pos + W] == 0)) o.tFS[last * W + j] = o.tFS[pos];
                    else o.tFS[last * W + j] = o.FS[last * W + j];
                }
            }
        }
    }
    for (int i = 1; i <= N; ++i)
        for (int j = 1; j <= M; ++j) o.FS[i * W + j] = o.tFS[i * W + j];
}

/* Scan_FSi + tridiagSolver + FSi_extract_solution + Update_FS,
 * Engine.py:338-361 (Sediment_Transport.cu:570-591). */
void sediment_i(Oracle& o, double t, double s_start, bool ketdinh)
{
    int W = o.W, N = o.N, M = o.M, SL = o.SL, tw = M + 3;
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 3; i <= N - 1; ++i) {
        for (int s = 0; s < o.moci[i]; ++s) {
            int first = o.daui[i * SL + s], last = o.cuoii[i * SL + s];
            if (last < first + 2) continue;
            bool br1, br2;
            wall_flags_v(o, i, first, last, br1, br2);
            for (int j = first + 1; j <= last - 1; ++j) fsi_matrix_cell(o, t, s_start, ketdinh, i, j, first, last, br1, br2);
            int sn = last - first - 2;
            o.SN[i * SL + s] = sn;
            int pos = i * tw + first + 1;
            tridiag(sn, &o.AA[pos], &o.BB[pos], &o.CC[pos], &o.DD[pos], &o.x[pos], &o.Ap[pos], &o.Bp[pos], &o.ep[pos]);
        }
    }
    for (int i = 3; i <= N - 1; ++i) {
        for (int s = 0; s < o.moci[i]; ++s) {
            int first = o.daui[i * SL + s], last = o.cuoii[i * SL + s];
            if (last < first + 2) continue;
            bool br1, br2;
            wall_flags_v(o, i, first, last, br1, br2);
            const double* x = &o.x[i * tw];
            for (int j = first + 1; j <= last - 1; ++j) {
                int pos = i * W + j;
                double xv = x[j];
                if (xv < 0) xv = o.P.NDnen;
                o.tFS[pos] = xv;
                if (j == first + 1) {
                    if ((br1) || (o.t_v[i * W + first] == 0)) o.tFS[i * W + first] = o.tFS[pos];
                    else o.tFS[i * W + first] = o.FS[i * W + first];
                }
                if (j == last - 1) {
                    if ((br2) || (o.t_v[i * W + last] == 0)) o.tFS[i * W + last] = o.tFS[pos];
                    else o.tFS[i * W + last] = o.FS[i * W + last];
                }
            }
        }
    }
    for (int i = 1; i <= N; ++i)
        for (int j = 1; j <= M; ++j) o.FS[i * W + j] = o.tFS[i * W + j];
}

/* Calculate_Qb, Sediment_Transport.cu:341-414.  D12: qb_mode REFERENCE keeps
 * the `!VTH[pos]` test, FIXED uses VTH != 0; the corner copies get the braces
 * the indentation shows. */
void calculate_qb(Oracle& o, int qb_mode)
{
    const hd_params& P = o.P;
    int W = o.W, N = o.N, M = o.M;
    double* Qbx = o.Qbx.data();
    double* Qby = o.Qby.data();
    for (int i = 2; i <= N; ++i)
        for (int j = 2; j <= M; ++j) {
            int pos = i * W + j;
            double Tob, Toee = P.Toe, Tx = 0;
            bool sel = (qb_mode == HD_QB_REFERENCE) ? (!o.VTH[pos]) : (o.VTH[pos] != 0);
            if (sel && (o.khouot[pos] == 0)) {
                Tob = P.ro * o.Fw[pos] * (o.VTH[pos] * o.VTH[pos]);
                if (o.H_moi[pos] > P.ghtoe) Toee = P.Toe * (1 + P.hstoe * ((o.H_moi[pos] - P.ghtoe)));
                if (Tob > Toee) Tx = Tob / Toee - 1;
                Qbx[pos] = 0.053 * hd_pow((P.Sx - 1) * P.g, 0.5) * hd_pow(P.dm, 1.5) * hd_pow(Tx, 2.1) * hd_pow(P.Dxr, -0.3)
                           * (o.t_u[pos] + o.t_u[pos - (planted(PL_QB_AVG_INDEX) ? 1 : W)]) * 0.5 / o.VTH[pos];
                Qby[pos] = 0.053 * hd_pow((P.Sx - 1) * P.g, 0.5) * hd_pow(P.dm, 1.5) * hd_pow(Tx, 2.1) * hd_pow(P.Dxr, -0.3)
                           * (o.t_v[pos] + o.t_v[pos - 1]) * 0.5 / o.VTH[pos];
            } else {
                Qbx[pos] = 0;
                Qby[pos] = 0;
            }
            if (i == 2) { Qbx[pos - W] = Qbx[pos]; Qby[pos - W] = Qby[pos]; }
            if (i == N) { Qbx[pos + W] = Qbx[pos]; Qby[pos + W] = Qby[pos]; }
            if (j == 2) {
                Qbx[pos - 1] = Qbx[pos]; Qby[pos - 1] = Qby[pos];
                if (i == 2) { Qbx[1 * W + 1] = Qbx[pos]; Qby[1 * W + 1] = Qby[pos]; }
                if (i == N) { Qbx[(N + 1) * W + 1] = Qbx[pos]; Qby[(N + 1) * W + 1] = Qby[pos]; }
            }
            if (j == M) {
                Qbx[pos + 1] = Qbx[pos]; Qby[pos + 1] = Qby[pos];
                if (i == 2) { Qbx[1 * W + M + 1] = Qbx[pos]; Qby[1 * W + M + 1] = Qby[pos]; }
                if (i == N) { Qbx[(N + 1) * W + M + 1] = Qbx[pos]; Qby[(N + 1) * W + M + 1] = Qby[pos]; }
            }
        }
}

/* _bed_load, Sediment_Transport.cu:416-506, via BedLoad :630-639 over rows
 * [sidx, eidx-1].  D13: `t_u[pos > 0]` reads t_u[1], as written. */
void bed_load(Oracle& o, bool ketdinh, int sidx, int eidx)
{
    const hd_params& P = o.P;
    int W = o.W, N = o.N, M = o.M, SL = o.SL;
    const double *FS = o.FS.data(), *Hm = o.H_moi.data(), *t_u = o.t_u.data(), *t_v = o.t_v.data(), *VTH = o.VTH.data(),
                 *Htdu = o.Htdu.data(), *Htdv = o.Htdv.data(), *Fw = o.Fw.data(), *Kx = o.Kx.data(), *Ky = o.Ky.data(),
                 *Qbx = o.Qbx.data(), *Qby = o.Qby.data();
    const int* kh = o.khouot.data();
    for (int i = sidx; i < eidx; ++i)
        for (int s = 0; s < o.moci[i]; ++s) {
            int first = o.daui[i * SL + s], last = o.cuoii[i * SL + s];
            if (last - first < 2) continue;
            for (int j = first; j <= last; ++j) {
                int pos = i * W + j;
                double p = 0, q = 0;
                if (kh[pos - W] == 1) {
                    if (t_u[pos] < 0) p = (-3 * Qbx[pos] + 4 * Qbx[pos + W] - Qbx[pos + 2 * W]) / P.dX2;
                } else {
                    if (kh[pos + W] == 1) {
                        if (t_u[1]) p = (3 * Qbx[pos] - 4 * Qbx[pos - W] + Qbx[pos - 2 * W]) / P.dX2;
                    } else p = (planted(PL_BED_DIV_SIGN) ? -1 : 1) * (Qbx[pos + W] - Qbx[pos - W]) / P.dX2;
                }
                if (kh[pos - 1] == 1) {
                    if (t_v[pos] < 0) q = (-3 * Qby[pos] + 4 * Qby[pos + 1] - Qby[pos + 2]) / P.dY2;
                } else {
                    if (kh[pos + 1] == 1) {
                        if (t_v[pos] > 0) q = (3 * Qby[pos] - 4 * Qby[pos - 1] + Qby[pos - 2]) / P.dY2;
                    } else q = (Qby[pos + 1] - Qby[pos - 1]) / P.dY2;
                }
                double tH = p + q;
                p = FS[pos + W] - FS[pos];
                if ((kh[pos + 2 * W] == 1 || kh[pos - W] == 1) && (i < N)) p = 0;
                q = FS[pos] - FS[pos - W];
                if ((i > 2) && (kh[pos - 2 * W] == 1 || kh[pos + W] == 1)) q = 0;
                p = 1 / P.dXbp * (Htdu[pos] * Kx[pos] * p - Htdu[pos - W] * Kx[pos - W] * q);
                tH = tH + p;
                p = FS[pos + 1] - FS[pos];
                if (((last < M) && (kh[pos - 1] == 1 || kh[pos + 2] == 1)) || (last == M && (kh[pos - 1] == 1))) p = 0;
                q = FS[pos] - FS[pos - 1];
                if (((first > 2) && (kh[pos - 2] == 1 || kh[pos + 1] == 1)) || (first == 2 && kh[pos + 1] == 1)) q = 0;
                p = 1 / (P.dYbp) * (Htdv[pos] * Ky[pos] * p - Htdv[pos - 1] * Ky[pos - 1] * q);
                tH = tH + p;
                double c = 18 * hd_log(12 * Hm[pos] / P.Ks);
                double Uf = std::sqrt(P.g) * std::fabs(VTH[pos]) / c;
                double wsm = P.wss * hd_pow((1 - FS[pos]), 4);
                double Zf = 0, S = 0;
                if (FS[pos] > 0.0) Zf = wsm / (0.4 * (Uf + 2 * wsm));
                if (ketdinh) S = source_chs(P, wsm, Zf, FS[pos], Fw[pos], VTH[pos], Hm[pos]);
                else S = source_nchs(P, wsm, Zf, P.Dxr, P.Ufr, Uf, Hm[pos], FS[pos]);
                o.dH[pos] += P.dT / (1 - P.Dorong) * (S + tH);
            }
        }
}

/* ------------------------------------------------------------------ step */

/* The two half-steps of Engine.py:144-366 after the t += dT/2 update. */
int halfstep(Oracle& o, int which, double t, const hd_step_opts& op, int* last_sidx, int* last_eidx)
{
    int N = o.N, M = o.M;
    bool channel = op.kenhhepd || op.kenhhepng;
    int bmode = op.boundary_mode;
    bool analytic = (bmode == HD_BND_ANALYTIC) || (bmode == HD_BND_AUTO && channel);
    int t1 = (int)(t / 3600);
    if (!analytic && (t1 + 1 >= o.total_time || t1 + 1 >= o.series_len)) {
        o.err = HD_ERANGE;
        o.msg = "boundary series too short for t";
        return o.err;
    }
    if (analytic) channel_boundaries(o, t);
    else update_boundary_value(o, t);
    int sidx, eidx;
    if (which == 0) {
        sidx = 2; eidx = M;
        if (channel && op.kenhhepd == 1) { sidx = 3; eidx = M - 1; }
        uz_solve(o, sidx, eidx);
        normalize(o, true);
        solve_v(o, 2, N);
        margins_v(o, 2, N);
        normalize(o, false);
        update_h_moi(o);
        reset_states_h(o);
    } else {
        sidx = 2; eidx = N;
        if (channel && op.kenhhepng == 1) { sidx = 3; eidx = N; }
        vz_solve(o, sidx, eidx);
        normalize(o, false);
        solve_u(o, 2, M);
        margins_u(o, 2, M);
        normalize(o, true);
        update_h_moi(o);
        reset_states_v(o);
    }
    update_uvz(o, op.kenhhepd, op.kenhhepng);
    find_limits(o);
    htuongdoi(o);
    if (t >= op.sediment_start) {
        find_vth(o);
        heso_k(o);
        if (which == 0) sediment_j(o, t, op.bed_change_start, op.ketdinh != 0);
        else { sediment_i(o, t, op.bed_change_start, op.ketdinh != 0); sidx = 3; eidx = N - 1; }
    }
    if (last_sidx) *last_sidx = sidx;
    if (last_eidx) *last_eidx = eidx;
    return o.err;
}

bool bed_trigger(double t, const hd_step_opts& op)
{
    if (op.bed_mode == HD_BED_EVERY_STEP) return t >= op.bed_change_start;
    double it = (double)(long long)t;
    if (t - it != 0) return false;
    if (op.bed_change_start == 0) return false;
    return std::fmod(it, op.bed_change_start) == 0;
}

int step(Oracle& o, double t, const hd_step_opts& op, double* t_out)
{
    if (!o.initialised) { o.err = HD_ESTATE; o.msg = "step before init"; return o.err; }
    t = t + o.P.dT * 0.5;
    int sidx = 0, eidx = 0;
    if (halfstep(o, 0, t, op, &sidx, &eidx)) return o.err;
    t = t + o.P.dT * 0.5;
    if (halfstep(o, 1, t, op, &sidx, &eidx)) return o.err;
    if (bed_trigger(t, op)) {
        calculate_qb(o, op.qb_mode);
        bed_load(o, op.ketdinh != 0, sidx, eidx);
    }
    if (t_out) *t_out = t;
    return o.err;
}

void resize_all(Oracle& o)
{
    int N = o.N, M = o.M, SL = o.SL;
    size_t F = (size_t)(N + 3) * (M + 3);
    for (const FieldRef& f : kFields) {
        if (!f.d) continue;
        (o.*(f.d)).assign(F, 0.0);
    }
    o.khouot.assign(F, 2);
    o.ubt.assign(M + 2, 0.0); o.ubp.assign(M + 2, 0.0);
    o.vbt.assign(N + 2, 0.0); o.vbd.assign(N + 2, 0.0);
    o.hi.assign(2 * (M + N + 6), 0.0);
    for (vd* s : {&o.bc_up, &o.bc_down, &o.bc_left, &o.bc_right, &o.CC_u, &o.CC_d, &o.CC_l, &o.CC_r})
        s->assign(o.series_len, 0.0);
    o.bienQ.assign(4, 0);
    o.boundary_type.assign(4 * SL, 0);
    o.daui.assign((N + 2) * SL, 0); o.cuoii.assign((N + 2) * SL, 0);
    o.dauj.assign((M + 2) * SL, 0); o.cuoij.assign((M + 2) * SL, 0);
    o.moci.assign(N + 2, 0); o.mocj.assign(M + 2, 0);
    int L = (N > M ? N : M);
    size_t S1 = (size_t)(L + 3) * (L + 3);
    for (vd* s : {&o.a1, &o.b1, &o.c1, &o.d1, &o.a2, &o.c2, &o.d2, &o.f1, &o.f2, &o.f3, &o.f5}) s->assign(S1, 0.0);
    size_t S2 = (size_t)(L + 3) * (2 * L + 4);
    for (vd* s : {&o.AA, &o.BB, &o.CC, &o.DD, &o.x, &o.Ap, &o.Bp, &o.ep}) s->assign(S2, 0.0);
    o.tmp.assign(F, 0.0);
    o.SN.assign((size_t)(L + 3) * SL, 0);
}

bool lookup(Oracle& o, const char* name, void** ptr, size_t* nbytes)
{
    for (const FieldRef& f : kFields) {
        if (std::strcmp(f.name, name) != 0) continue;
        if (f.d) { vd& v = o.*(f.d); *ptr = v.data(); *nbytes = v.size() * sizeof(double); }
        else { vi& v = o.*(f.iv); *ptr = v.data(); *nbytes = v.size() * sizeof(int); }
        return true;
    }
    return false;
}

} // namespace

/* ------------------------------------------------------------------ C ABI */

extern "C" {

void* or_create(const hd_params* p, int N, int M, int series_len, int total_time)
{
    if (!p || N < 4 || M < 4 || p->segment_limit < 1) return nullptr;
    Oracle* o = new Oracle();
    o->P = *p;
    derive_params(o->P);
    o->N = N; o->M = M; o->SL = p->segment_limit; o->W = M + 3;
    o->series_len = series_len; o->total_time = total_time;
    resize_all(*o);
    return o;
}

void or_destroy(void* h) { delete (Oracle*)h; }

const char* or_last_error(void* h) { return ((Oracle*)h)->msg.c_str(); }

int or_get_params(void* h, hd_params* out) { *out = ((Oracle*)h)->P; return 0; }

int or_set(void* h, const char* name, const void* src, size_t nbytes)
{
    Oracle& o = *(Oracle*)h;
    void* ptr; size_t nb;
    if (!lookup(o, name, &ptr, &nb)) return HD_ENAME;
    if (nb != nbytes) return HD_ESHAPE;
    std::memcpy(ptr, src, nbytes);
    return 0;
}

int or_get(void* h, const char* name, void* dst, size_t nbytes)
{
    Oracle& o = *(Oracle*)h;
    void* ptr; size_t nb;
    if (!lookup(o, name, &ptr, &nb)) return HD_ENAME;
    if (nb != nbytes) return HD_ESHAPE;
    std::memcpy(dst, ptr, nbytes);
    return 0;
}

int or_get_hmax(void* h, double* out)
{
    Oracle& o = *(Oracle*)h;
    out[0] = o.hmax_u; out[1] = o.hmax_d; out[2] = o.hmax_l; out[3] = o.hmax_r;
    return 0;
}

/* rep.py:126-139 */
int or_init(void* h)
{
    Oracle& o = *(Oracle*)h;
    o.err = 0;
    onetime_init(o);
    find_limits(o);
    htuongdoi(o);
    preprocess(o);
    o.initialised = (o.err == 0);
    return o.err;
}

int or_step(void* h, double t, const hd_step_opts* op, double* t_out)
{
    Oracle& o = *(Oracle*)h;
    return step(o, t, *op, t_out);
}

int or_run(void* h, double t, int nsteps, const hd_step_opts* op, double* t_out)
{
    Oracle& o = *(Oracle*)h;
    for (int k = 0; k < nsteps; ++k)
        if (step(o, t, *op, &t)) return o.err;
    if (t_out) *t_out = t;
    return 0;
}

int or_halfstep(void* h, int which, double t, const hd_step_opts* op)
{
    Oracle& o = *(Oracle*)h;
    return halfstep(o, which, t, *op, nullptr, nullptr);
}

int or_bed_change(void* h, double t, const hd_step_opts* op, int sidx, int eidx)
{
    Oracle& o = *(Oracle*)h;
    (void)t;
    calculate_qb(o, op->qb_mode);
    bed_load(o, op->ketdinh != 0, sidx, eidx);
    return o.err;
}

/* Segment finder alone (support_funcs.cu:230-335) on the current khouot. */
int or_find_limits(void* h)
{
    Oracle& o = *(Oracle*)h;
    o.err = 0;
    find_limits(o);
    return o.err;
}

long long or_wet_count(void* h)
{
    Oracle& o = *(Oracle*)h;
    long long n = 0;
    for (int k : o.khouot) n += (k == 0);
    return n;
}

/* Thomas on one system of size sn+1 (UVZSolver_multithread.cu:203-228). */
int or_thomas(int sn, const double* A, const double* B, const double* C, const double* D, double* x)
{
    std::vector<double> Ap(sn + 1), Bp(sn + 1), ep(sn + 1);
    tridiag(sn, A, B, C, D, x, Ap.data(), Bp.data(), ep.data());
    return 0;
}

/* Scan of one line mask (support_funcs.cu:252-279) for exhaustive tests. */
int or_scan_line(const int* kh, int lim, int SL, int* dau, int* cuoi)
{
    int moc = 0;
    int r = scan_line_q(SL, 0, lim, kh, 1, dau, cuoi, &moc);
    return r < 0 ? HD_ESEGOVERFLOW : moc;
}

/* Threads used by the OpenMP loops over independent lines / cells
 * (results do not depend on it). */
int or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* One restated kernel on the current state, for the tests that pin each
 * against closed forms and for the per-kernel GPU parity tests.  Arguments by
 * kernel (a, b, c integers, t a time):
 *   uz_preindex / vz_preindex (a = i, b = j): one cell, into f1..f5, a2, c2, d2
 *   uz_solve / vz_solve (a = first line, b = last line): preindex, abcd,
 *       matrix_coeff, tridiag and extraction of those lines
 *   u_solver / v_solver (a = i, b = j): one explicit cell into t_u / t_v
 *   solve_u / solve_v, margins_u / margins_v (a, b = line range)
 *   normalize (a = 1: t_u, 0: t_v), update_h_moi, reset_states_h,
 *   reset_states_v, update_uvz (a = kenhhepd, b = kenhhepng), find_limits,
 *   htuongdoi, find_vth, heso_k, calculate_qb (a = qb_mode),
 *   bed_load (a = ketdinh, b = sidx, c = eidx),
 *   sediment_j / sediment_i (a = ketdinh, t; the scan, solve, extraction and
 *       update of the FS sweep, bed-change start taken from c as seconds),
 *   boundary (t: Update_Boundary_Value), channel_boundary (t),
 *   onetime_init, preprocess. */
int or_kernel(void* h, const char* name, int a, int b, int c, double t)
{
    Oracle& o = *(Oracle*)h;
    o.err = 0;
    std::string n(name);
    if (n == "onetime_init") onetime_init(o);
    else if (n == "preprocess") preprocess(o);
    else if (n == "uz_preindex") uz_preindex(o, a, b);
    else if (n == "vz_preindex") vz_preindex(o, a, b);
    else if (n == "uz_solve") uz_solve(o, a, b);
    else if (n == "vz_solve") vz_solve(o, a, b);
    else if (n == "u_solver") u_solver(o, a, b);
    else if (n == "v_solver") v_solver(o, a, b);
    else if (n == "solve_u") solve_u(o, a, b);
    else if (n == "solve_v") solve_v(o, a, b);
    else if (n == "margins_u") margins_u(o, a, b);
    else if (n == "margins_v") margins_v(o, a, b);
    else if (n == "normalize") normalize(o, a != 0);
    else if (n == "update_h_moi") update_h_moi(o);
    else if (n == "reset_states_h") reset_states_h(o);
    else if (n == "reset_states_v") reset_states_v(o);
    else if (n == "update_uvz") update_uvz(o, a, b);
    else if (n == "find_limits") find_limits(o);
    else if (n == "htuongdoi") htuongdoi(o);
    else if (n == "find_vth") find_vth(o);
    else if (n == "heso_k") heso_k(o);
    else if (n == "calculate_qb") calculate_qb(o, a);
    else if (n == "bed_load") bed_load(o, a != 0, b, c);
    else if (n == "sediment_j") sediment_j(o, t, (double)c, a != 0);
    else if (n == "sediment_i") sediment_i(o, t, (double)c, a != 0);
    else if (n == "boundary") update_boundary_value(o, t);
    else if (n == "channel_boundary") channel_boundaries(o, t);
    else { o.msg = "unknown kernel " + n; return HD_ENAME; }
    return o.err;
}

/* source_chs / source_nchs (Sediment_Transport.cu:12-60) with the case's
 * parameters, for closed-form checks. */
double or_source_chs(void* h, double wsm, double Zf, double fs, double fw, double vth, double h_moi)
{
    return source_chs(((Oracle*)h)->P, wsm, Zf, fs, fw, vth, h_moi);
}

double or_source_nchs(void* h, double wsm, double Zf, double Uf, double h_moi, double fs)
{
    const hd_params& P = ((Oracle*)h)->P;
    return source_nchs(P, wsm, Zf, P.Dxr, P.Ufr, Uf, h_moi, fs);
}

/* Plant mistake `code` (enum Plant) in the restated kernels, 0 to clear;
 * returns the previous code.  Only the tests that show a pin catches its
 * mistake set it. */
int or_plant(int code)
{
    int prev = g_plant;
    g_plant = code;
    return prev;
}

/* hd_math entry points, for accuracy tests against libm. */
double or_hd_pow(double x, double y) { return hd_pow(x, y); }
double or_hd_log(double x) { return hd_log(x); }
double or_hd_exp(double x) { return hd_exp(x); }
double or_hd_cos(double x) { return hd_cos(x); }
double or_hd_sin(double x) { return hd_sin(x); }

} // extern "C"
