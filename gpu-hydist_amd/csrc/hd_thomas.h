/*
 * hd_thomas.h -- the speculative parallel Thomas solve shared by the ADI and
 * FS run kernels (one wave of 64 lanes per tridiagonal system, rows in LDS).
 */
#ifndef HD_THOMAS_H
#define HD_THOMAS_H

#include <hip/hip_runtime.h>

namespace hd_spec {

constexpr int kSpecWave = 64;

/* ------------------------------------------- speculative parallel Thomas
 *
 * The forward elimination of tridiag (UVZSolver_multithread.cu:203-228),
 *   e_i = A_i ap_{i-1} + B_i,  ap_i = -C_i / e_i,  bp_i = (D_i - A_i bp_{i-1}) / e_i,
 * is a serial chain of a multiply, an add and a correctly rounded division per
 * row.  In exact arithmetic (ap, bp) = (p / q, r / q) where (p, q, r) follows
 * the linear recurrence v_i = M_i v_{i-1},
 *   M_i = [[0, -C_i, 0], [A_i, B_i, 0], [0, D_i, -A_i]],
 * so each lane composes the matrices of its own rows, an inclusive scan over
 * the wave (DPP row shifts and broadcasts) composes the prefixes, and every
 * lane obtains an estimate, good to
 * a few ulps, of the state after each earlier lane's rows.  The recurrence
 * contracts on these systems (each continuity / momentum pair reduces to a
 * diagonally dominant system in z), so two rounded trajectories that start a
 * few ulps apart become bit-identical after a few tens of rows.  A lane
 * therefore runs the exact recurrence -- the same IEEE operations in the same
 * order as the serial solve -- from the estimate of the state kSpecWF rows
 * (rounded up to whole lanes) before its own rows, through its own rows, and
 * the result is verified exactly: the state a lane had on entering its rows
 * must equal, bit for bit, the state its predecessor produced at its last
 * row.  Lane 0 starts from the exact state; by induction every lane that
 * passes is exact, and the first lane that fails is recomputed serially from
 * its verified predecessor, after which its successor is tested again.  The
 * result equals the serial solve's bits in every case; the speculation only
 * shortens the critical path from the run's length to about kSpecWF + L rows
 * (L = rows per lane).  Back substitution, x_i = bp_i + ap_i x_{i+1}, is an
 * affine chain and is treated the same way with 2-entry affine maps.
 *
 * Repairs (spec_forward<true>, the ADI sweeps): while a flow starts from
 * rest, bp decays ahead of the waves as fast as the estimate's relative
 * error and never merges, so long stretches of lanes fail with their ap
 * already exact.  Then every row's divisor is known, and a failing lane
 * redoes only its bp chain, from registers, with the division in
 * Markstein's reciprocal form (bp_known_e) -- the same rounding as the
 * division; spec_forward<false> (the FS sweeps, which need the registers
 * for occupancy) keeps the plain repair. */
constexpr int kSpecWF = 16;           /* forward warm-up rows after the estimate */
constexpr int kSpecWB = 32;           /* back-substitution warm-up rows */
constexpr int kSpecSerialF = 32;      /* up to this many rows the forward sweep stays on one lane */
constexpr int kSpecSerialB = 96;      /* likewise for back substitution */
constexpr int kSpecMaxL = 8;          /* rows per lane of one speculative sweep (chunks <= 512 rows) */

struct Mx {
    double a, b, c, d, e, f, g; /* [[a b 0] [c d 0] [e f g]] */
};

/* M_i X for the row matrix M_i of (A, B, C, D) */
__device__ __forceinline__ Mx mx_row(double A, double B, double C, double D, const Mx& X)
{
    Mx R;
    R.a = -C * X.c;
    R.b = -C * X.d;
    R.c = fma(A, X.a, B * X.c);
    R.d = fma(A, X.b, B * X.d);
    R.e = fma(D, X.c, -A * X.e);
    R.f = fma(D, X.d, -A * X.f);
    R.g = -A * X.g;
    return R;
}

/* X Y */
__device__ __forceinline__ Mx mx_mul(const Mx& X, const Mx& Y)
{
    Mx R;
    R.a = fma(X.a, Y.a, X.b * Y.c);
    R.b = fma(X.a, Y.b, X.b * Y.d);
    R.c = fma(X.c, Y.a, X.d * Y.c);
    R.d = fma(X.c, Y.b, X.d * Y.d);
    R.e = fma(X.e, Y.a, fma(X.f, Y.c, X.g * Y.e));
    R.f = fma(X.e, Y.b, fma(X.f, Y.d, X.g * Y.f));
    R.g = X.g * Y.g;
    return R;
}

/* Scale by a power of two (exact) so the largest entry lies in [0.5, 1):
 * only ratios of the composed state matter. */
__device__ __forceinline__ void mx_norm(Mx& X)
{
    double m = fmax(fmax(fmax(fabs(X.a), fabs(X.b)), fmax(fabs(X.c), fabs(X.d))),
                    fmax(fmax(fabs(X.e), fabs(X.f)), fabs(X.g)));
    if (m > 0 && m <= 1.7976931348623157e308) {
        int ex;
        frexp(m, &ex);
        double s = ldexp(1.0, -ex);
        X.a *= s; X.b *= s; X.c *= s; X.d *= s; X.e *= s; X.f *= s; X.g *= s;
    }
}

/* A 64-bit value moved across lanes by a DPP mode (row shifts within rows
 * of 16 lanes, row broadcasts to later rows): a lane the mode gives no source,
 * or whose row ROWMASK leaves out, gets `old`. */
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double old, double src)
{
    const long long o = __double_as_longlong(old), v = __double_as_longlong(src);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)v, CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(v >> 32), CTRL, ROWMASK, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

/* One step of the wave's inclusive prefix scan of row matrices: X := X Y
 * with Y the matrix of the source lane, the identity where there is none. */
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void mx_scan_step(Mx& X)
{
    Mx Y;
    Y.a = dpp_f64<CTRL, ROWMASK>(1.0, X.a); Y.b = dpp_f64<CTRL, ROWMASK>(0.0, X.b);
    Y.c = dpp_f64<CTRL, ROWMASK>(0.0, X.c); Y.d = dpp_f64<CTRL, ROWMASK>(1.0, X.d);
    Y.e = dpp_f64<CTRL, ROWMASK>(0.0, X.e); Y.f = dpp_f64<CTRL, ROWMASK>(0.0, X.f);
    Y.g = dpp_f64<CTRL, ROWMASK>(1.0, X.g);
    X = mx_mul(X, Y);
    mx_norm(X);
}

/* The same for the affine maps x -> al x + be of back substitution: the
 * source lane's map is applied first. */
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void aff_scan_step(double& al, double& be)
{
    const double a2 = dpp_f64<CTRL, ROWMASK>(1.0, al), b2 = dpp_f64<CTRL, ROWMASK>(0.0, be);
    be = fma(al, b2, be);
    al = al * a2;
}

/* v of lane `l` (wave-uniform l) through the scalar unit */
__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ bool same_bits(double x, double y)
{
    return __double_as_longlong(x) == __double_as_longlong(y);
}

/* One exact forward row, in the operation order of tridiag. */
__device__ __forceinline__ void thomas_row(const double* sA, const double* sB, const double* sC, const double* sD,
                                           int i, double& ap, double& bp)
{
    double A = sA[i], B = sB[i], C = sC[i], D = sD[i];
    double e = A * ap + B;
    double nap = -C / e;
    double nbp = (D - (A * bp)) / e;
    ap = nap;
    bp = nbp;
}

/* bp of one forward row when its divisor e and y = RN(1/e) are known
 * (Markstein's correction: q0 = RN(n y) is within an ulp of n / e, the
 * remainder n - q0 e is exact by FMA, and RN(q0 + r y) is then RN(n / e)):
 * the bits of (D - A bp) / e with a multiply and two FMAs instead of a
 * division, valid where no intermediate underflows or overflows (|n| >=
 * 2^-960, |q0| in [2^-1000, 2^1000]; e and y are checked by the caller).
 * Branch-free: `off` is set when a row leaves that range, and the caller
 * then redoes the rows with divisions. */
__device__ __forceinline__ double bp_known_e(double A, double D, double bp, double e, double y, bool& off)
{
    const double n = D - (A * bp);
    const double q0 = n * y;
    const double an = fabs(n), aq = fabs(q0);
    off = off || !(an >= 0x1p-960 && an <= 0x1p1000 && aq >= 0x1p-1000 && aq <= 0x1p1000);
    const double r = fma(-q0, e, n);
    return fma(r, y, q0);
}

/* Forward elimination of LDS rows g0 .. g1 (chunk-relative), entering with
 * the exact state (ap, bp) of row g0 - 1, which is left as the state of row
 * g1.  The factors ap / bp of the rows replace C / D (after every lane has
 * read them: speculating lanes keep their own rows' factors in registers
 * until the sweep is verified).  At most 64 * kSpecMaxL rows; one wave,
 * uniform control; ends with a barrier. */
template <bool kFastRepair>
__device__ __forceinline__ void spec_forward(const double* sA, double* sB, double* sC, double* sD, int g0,
                                             int g1, double& ap, double& bp, int minL = 1, int* fails = nullptr)
{
    const int lane = threadIdx.x;
    const int n = g1 - g0 + 1;
    if (n <= 0) return;
    if (n <= kSpecSerialF) {
        /* short sweep: lane 0 alone, factors written in place */
        double ca = ap, cb = bp;
        if (lane == 0) {
#pragma unroll 4
            for (int i = g0; i <= g1; ++i) {
                thomas_row(sA, sB, sC, sD, i, ca, cb);
                sC[i] = ca;
                sD[i] = cb;
            }
        }
        ap = __shfl(ca, 0, kSpecWave);
        bp = __shfl(cb, 0, kSpecWave);
        __syncthreads();
        return;
    }
    /* rows per lane, <= kSpecMaxL: at least minL (more rows per lane, fewer
     * lanes, a shorter chain of lane-by-lane repairs when speculation fails;
     * the result is the same for any L) */
    const int L0 = (n + kSpecWave - 1) / kSpecWave;
    const int L = L0 < minL ? minL : L0;
    const int sk = g0 + lane * L;
    const int ek = sk + L <= g1 + 1 ? sk + L : g1 + 1;
    const bool act = sk <= g1;
    const int m = (kSpecWF + L - 1) / L;
    const int j = lane - m - 1;
    /* prefix estimates */
    Mx X{1, 0, 0, 1, 0, 0, 1};
    for (int i = sk; i < ek; ++i) {
        X = mx_row(sA[i], sB[i], sC[i], sD[i], X);
        if (((i - sk) & 3) == 3) mx_norm(X);
    }
    mx_norm(X);
    /* inclusive scan over the wave with DPP moves: shifts by 1, 2, 4, 8
     * within rows of 16 lanes, then the end of row r to row r + 1 (rows 1
     * and 3), then the end of row 1 to rows 2 and 3; the estimates need not
     * match any particular rounding */
    mx_scan_step<0x111, 0xF>(X);
    mx_scan_step<0x112, 0xF>(X);
    mx_scan_step<0x114, 0xF>(X);
    mx_scan_step<0x118, 0xF>(X);
    mx_scan_step<0x142, 0xA>(X);
    mx_scan_step<0x143, 0xC>(X);
    double ea, eb;
    {
        double p = fma(X.a, ap, X.b), q = fma(X.c, ap, X.d), r = fma(X.e, ap, fma(X.g, bp, X.f));
        ea = __shfl(p / q, j < 0 ? 0 : j, kSpecWave);
        eb = __shfl(r / q, j < 0 ? 0 : j, kSpecWave);
    }
    const bool exact = j < 0;
    double ca = exact ? ap : ea, cb = exact ? bp : eb;
    int i0 = exact ? g0 : g0 + (j + 1) * L;
    if (!act) i0 = sk;
    /* warm-up rows, then the lane's own rows */
    for (int i = i0; i < sk; ++i) thomas_row(sA, sB, sC, sD, i, ca, cb);
    const double wa = ca, wb = cb;
    double ra[kSpecMaxL], rb[kSpecMaxL];
#pragma unroll
    for (int q = 0; q < kSpecMaxL; ++q) {
        if (sk + q < ek) {
            thomas_row(sA, sB, sC, sD, sk + q, ca, cb);
            ra[q] = ca;
            rb[q] = cb;
        }
    }
    /* exact verification against the predecessor's last row */
    {
        double pa = __shfl_up(ca, 1, kSpecWave), pb = __shfl_up(cb, 1, kSpecWave);
        bool ok = exact || !act || (same_bits(wa, pa) && same_bits(wb, pb));
        unsigned long long bad = __ballot(!ok);
        if (bad && fails && lane == 0) atomicAdd(fails, 1);
        if constexpr (kFastRepair) {
            /* Where the flow starts from rest every failure has the lane's
             * ap bits right (ap depends on A, B, C only and merges; bp decays
             * ahead of the waves as fast as the estimate's error), so the
             * lane's rows' divisors e are exact: every lane takes A, D, e
             * and 1/e of its rows into registers once a failure is seen, and
             * the repair of such a lane runs bp's chain in the reciprocal
             * form (bp_known_e) with no LDS access; the predecessor's state
             * moves through the scalar unit. */
            double rA[kSpecMaxL], rD[kSpecMaxL], re[kSpecMaxL], ry[kSpecMaxL];
            bool mk = false;
            if (bad) {
                mk = true;
#pragma unroll
                for (int q = 0; q < kSpecMaxL; ++q) {
                    rA[q] = 0.0; rD[q] = 0.0; re[q] = 1.0; ry[q] = 1.0;
                    if (sk + q < ek) {
                        rA[q] = sA[sk + q];
                        rD[q] = sD[sk + q];
                        re[q] = rA[q] * (q == 0 ? wa : ra[q - 1]) + sB[sk + q];
                        ry[q] = 1.0 / re[q];
                        const double ae = fabs(re[q]);
                        mk = mk && ae >= 0x1p-1000 && ae <= 0x1p1000;
                    }
                }
            }
            while (bad) {
                const int f = __ffsll((long long)bad) - 1; /* >= 1: lane 0 starts exact */
                double fa = readlane_f64(ca, f - 1), fb = readlane_f64(cb, f - 1);
                if (lane == f) {
                    if (mk && same_bits(wa, fa)) {
                        /* ap of the lane's rows is exact already: bp only */
                        bool off = false;
                        cb = fb;
#pragma unroll
                        for (int q = 0; q < kSpecMaxL; ++q) {
                            if (sk + q < ek) {
                                cb = bp_known_e(rA[q], rD[q], cb, re[q], ry[q], off);
                                rb[q] = cb;
                            }
                        }
                        if (off) { /* a row outside the reciprocal form's range: divide */
                            cb = fb;
#pragma unroll
                            for (int q = 0; q < kSpecMaxL; ++q) {
                                if (sk + q < ek) {
                                    cb = (rD[q] - (rA[q] * cb)) / re[q];
                                    rb[q] = cb;
                                }
                            }
                        }
                    } else {
                        ca = fa;
                        cb = fb;
#pragma unroll
                        for (int q = 0; q < kSpecMaxL; ++q) {
                            if (sk + q < ek) {
                                thomas_row(sA, sB, sC, sD, sk + q, ca, cb);
                                ra[q] = ca;
                                rb[q] = cb;
                            }
                        }
                    }
                }
                double na = readlane_f64(ca, f), nb = readlane_f64(cb, f);
                if (lane == f) ok = true;
                if (lane == f + 1) ok = exact || !act || (same_bits(wa, na) && same_bits(wb, nb));
                bad = __ballot(!ok);
            }
        } else {
            while (bad) {
                const int f = __ffsll((long long)bad) - 1; /* >= 1: lane 0 starts exact */
                double fa = __shfl(ca, f - 1, kSpecWave), fb = __shfl(cb, f - 1, kSpecWave);
                if (lane == f) {
                    ca = fa;
                    cb = fb;
#pragma unroll
                    for (int q = 0; q < kSpecMaxL; ++q) {
                        if (sk + q < ek) {
                            thomas_row(sA, sB, sC, sD, sk + q, ca, cb);
                            ra[q] = ca;
                            rb[q] = cb;
                        }
                    }
                }
                double na = __shfl(ca, f, kSpecWave), nb = __shfl(cb, f, kSpecWave);
                if (lane == f) ok = true;
                if (lane == f + 1) ok = exact || !act || (same_bits(wa, na) && same_bits(wb, nb));
                bad = __ballot(!ok);
            }
        }
    }
    const int klast = (n - 1) / L;
    ap = __shfl(ca, klast, kSpecWave);
    bp = __shfl(cb, klast, kSpecWave);
    __syncthreads(); /* every lane has read C / D */
#pragma unroll
    for (int q = 0; q < kSpecMaxL; ++q) {
        if (sk + q < ek) {
            sC[sk + q] = ra[q];
            sD[sk + q] = rb[q];
        }
    }
    __syncthreads();
}

/* Back substitution over LDS rows hi down to 0 (chunk-relative) given x of
 * row hi + 1; x_i goes to sX[i].  Returns x of row 0 (x_in when hi < 0). */
__device__ __forceinline__ double spec_backward(const double* sAp, const double* sBp, double* sX, int hi, double x_in)
{
    const int lane = threadIdx.x;
    const int n = hi + 1;
    if (n <= 0) return x_in;
    const int L = n <= kSpecSerialB ? n : (n + kSpecWave - 1) / kSpecWave;
    const int top = hi - lane * L;
    const int bot = top - L + 1 >= 0 ? top - L + 1 : 0;
    const bool act = top >= 0;
    const int m = (kSpecWB + L - 1) / L;
    const int j = lane - m - 1;
    double cx = x_in;
    int i0 = hi;
    if (n > kSpecSerialB) {
        double al = 1, be = 0;
        for (int i = top; i >= bot; --i) {
            be = fma(sAp[i], be, sBp[i]);
            al = sAp[i] * al;
        }
        /* inclusive scan of the affine maps, DPP moves as in spec_forward */
        aff_scan_step<0x111, 0xF>(al, be);
        aff_scan_step<0x112, 0xF>(al, be);
        aff_scan_step<0x114, 0xF>(al, be);
        aff_scan_step<0x118, 0xF>(al, be);
        aff_scan_step<0x142, 0xA>(al, be);
        aff_scan_step<0x143, 0xC>(al, be);
        double est = fma(al, x_in, be);
        est = __shfl(est, j < 0 ? 0 : j, kSpecWave);
        if (j >= 0) { cx = est; i0 = hi - (j + 1) * L; }
    }
    const bool exact = j < 0;
    if (!act) i0 = top;
    for (int i = i0; i > top; --i) cx = sBp[i] + (sAp[i] * cx);
    const double w = cx;
#pragma unroll 8
    for (int i = (top < i0 ? top : i0); i >= bot; --i) {
        cx = sBp[i] + (sAp[i] * cx);
        sX[i] = cx;
    }
    if (n > kSpecSerialB) {
        double px = __shfl_up(cx, 1, kSpecWave);
        bool ok = exact || !act || same_bits(w, px);
        unsigned long long bad = __ballot(!ok);
        while (bad) {
            const int f = __ffsll((long long)bad) - 1;
            __syncthreads();
            if (lane == f) {
                cx = sX[top + 1];
                for (int i = top; i >= bot; --i) {
                    cx = sBp[i] + (sAp[i] * cx);
                    sX[i] = cx;
                }
            }
            double nx = __shfl(cx, f, kSpecWave);
            if (lane == f) ok = true;
            if (lane == f + 1) ok = exact || !act || same_bits(w, nx);
            bad = __ballot(!ok);
        }
    }
    return __shfl(cx, (n - 1) / L, kSpecWave);
}

} // namespace hd_spec

#endif
