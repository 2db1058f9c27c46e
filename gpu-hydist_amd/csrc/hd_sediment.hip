/*
 * hd_sediment.hip -- suspended-sediment sweep and bed change
 * (Sediment_Transport.cu), gfx950.
 *
 * Per half-step with t >= sediment_start: VTH + hesoK (one thread per cell),
 * the concentration rows of the sweep direction (one thread per interior cell
 * of each wet run), one lane per line for the Thomas solve with the
 * extraction into tFS, then FS = tFS.  Bed change: Calculate_Qb then
 * _bed_load, one thread per cell.  Formulas follow the cited reference lines
 * in the same operation order (contraction off).
 */
#include "hd_internal.h"
#include "hd_thomas.h"

namespace {

/* Cell (i, j) of a thread over the context's active 64-column row tiles
 * (tiles holding or bordering a cell with khouot != 2): outside them the
 * sediment kernels that run by tiles have reached a fixed point once they
 * have run over the whole grid, which the caller arranges. */
__device__ __forceinline__ bool tile_cell(const HdDevG& d, int& i, int& j)
{
    const int t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (t >= d.ntiles[0]) return false;
    const int tile = d.tiles[t];
    const int ncb = (d.M + 3 + 63) >> 6;
    i = tile / ncb;
    j = (tile % ncb) * 64 + (threadIdx.x & 63);
    return j < d.M + 3 && i < d.N + 3;
}

/* source_chs, Sediment_Transport.cu:12-49. */
__device__ double source_chs(const hd_params& P, double wsm, double Zf, double fs, double fw, double vth, double h_moi)
{
    double toee = P.Toe, todd = P.Tod, S = 0;
    double tob = P.ro * fw * (vth * vth);
    if (h_moi > P.ghtoe) toee = P.Toe * (1 + P.hstoe * ((h_moi - P.ghtoe)));
    if (tob > toee) S = P.Mbochat / P.ros * (tob - toee) / toee;
    else if (tob < todd) {
        double Pd = 1 - (tob / todd);
        double beta = 1 + (Zf / (1.25 + 4.75 * hd_pow(Pd, 2.5)));
        double Cb = beta * fs;
        S = -wsm * Cb * Pd;
    }
    return S;
}

/* source_nchs, Sediment_Transport.cu:52-60. */
__device__ double source_nchs(const hd_params& P, double wsm, double Zf, double Uf, double h_moi, double fs)
{
    double gamac = 0.434 + 5.975 * Zf + 2.888 * Zf * Zf;
    double Tx = fmax(0.0, (Uf * Uf - P.Ufr * P.Ufr) / (P.Ufr * P.Ufr));
    double Cac = 0.015 * P.dm * hd_pow(Tx, 1.5) * hd_pow(P.Dxr, (-0.3)) / (0.05 * h_moi);
    return wsm * (Cac - gamac * fs);
}

/* Find_VTH (Sediment_Transport.cu:541-568) and hesoK (:508-539). */
__global__ void __launch_bounds__(256) k_vth_hesok(const HdDev* __restrict__ dv, int tiled)
{
    const HdDevG& d = dev_view(dv);
    int j, i;
    if (tiled) {
        if (!tile_cell(d, i, j)) return;
    } else {
        j = blockIdx.x * blockDim.x + threadIdx.x;
        i = blockIdx.y;
    }
    int N = d.N, M = d.M, W = d.W;
    const hd_params& P = d.P;
    if (j > M || i > N) return;
    int pos = i * W + j;
    if (i >= 2 && j >= 2) {
        if ((d.htaiz[pos] > P.NANGDAY) && (d.H_moi[pos] > P.H_TINH)) {
            double ut = (d.t_u[pos] + d.t_u[pos - W]) * 0.5;
            double vt = (d.t_v[pos] + d.t_v[pos - 1]) * 0.5;
            d.VTH[pos] = sqrt(ut * ut + vt * vt);
        } else d.VTH[pos] = 0;
    }
    if (i >= 1 && j >= 1) {
        double Cz;
        if (d.Htdu[pos] > 0) {
            Cz = 18 * hd_log(12 * d.Htdu[pos] / P.Ks);
            double kx = 5.93 * sqrt(P.g) * (d.h[pos - 1] + d.h[pos]) * 0.25 * fabs(d.t_u[pos]) / Cz;
            d.Kx[pos] = fmin(100.0, fmax(5.0, kx));
        }
        if (d.Htdv[pos] > 0) {
            Cz = 18 * hd_log(12 * d.Htdv[pos] / P.Ks);
            double ky = 5.93 * sqrt(P.g) * (d.h[pos - W] + d.h[pos]) * 0.25 * fabs(d.t_v[pos]) / Cz;
            d.Ky[pos] = fmin(100.0, fmax(5.0, ky));
        }
    }
}

__device__ __forceinline__ void wall_flags2(const HdDevG& d, bool isJ, int line, int first, int last, bool& br1, bool& br2)
{
    int W = d.W;
    const double* h = d.h;
    double ND = d.P.NANGDAY;
    if (isJ) { /* locate_segment_u, column line */
        br1 = (first > 2) || ((first == 2) && ((h[1 * W + line] + h[1 * W + line - 1]) * 0.5 == ND));
        br2 = (last < d.N) || ((last == d.N) && ((h[d.N * W + line] + h[d.N * W + line - 1]) * 0.5 == ND));
    } else {   /* locate_segment_v, row line */
        br1 = (first > 2) || ((first == 2) && ((h[line * W + first - 1] + h[(line - 1) * W + first - 1]) * 0.5 == ND));
        br2 = (last < d.M) || ((last == d.M) && ((h[line * W + last] + h[(line - 1) * W + last]) * 0.5 == ND));
    }
}

struct FsRow {
    double A, B, C, D;
};

/* One concentration row at cell (i, j) of run [first, last]:
 * _FSj_calculate__mactrix_coeff (Sediment_Transport.cu:205-302, column lines,
 * gamav before Zf and && at the last row, as written) or
 * _FSi_calculate__mactrix_coeff (:62-165, row lines). */
template <bool isJ>
__device__ FsRow fs_row(const HdDevG& d, double t, double s_start, int ketdinh, int i, int j, int first, int last,
                        bool br1, bool br2)
{
    const hd_params& P = d.P;
    int W = d.W;
    const double *FS = d.FS, *Hm = d.H_moi, *t_u = d.t_u, *t_v = d.t_v, *VTH = d.VTH, *Htdu = d.Htdu, *Htdv = d.Htdv,
                 *Fw = d.Fw, *Kx = d.Kx, *Ky = d.Ky;
        int pos = i * W + j;
        double c = 18 * hd_log(12 * Hm[pos] / P.Ks);
        double Uf = sqrt(P.g) * fabs(VTH[pos]) / c;
        double wsm = P.wss * hd_pow((1 - FS[pos]), 4);
        double Zf = 0, S = 0, gamav;
        if (isJ) gamav = 0.98 - 0.198 * Zf + 0.032 * Zf * Zf;
        if (FS[pos] > 0.00001) Zf = wsm / (0.4 * (Uf + 2 * wsm));
        if (ketdinh) S = source_chs(P, wsm, Zf, FS[pos], Fw[pos], VTH[pos], Hm[pos]);
        else S = source_nchs(P, wsm, Zf, Uf, Hm[pos], FS[pos]);
        if (t < s_start) S = 0;
        if (!isJ) gamav = 0.98 - 0.198 * Zf + 0.032 * Zf * Zf;
        double A, B, C, D, p = 0, q = 0;
        if (isJ) {
            A = -gamav * 0.5 * (t_u[pos] + t_u[pos - W]) / P.dX2 - Htdu[pos - W] * Kx[pos - W] / (Hm[pos] * (P.dXbp));
            C = gamav * 0.5 * (t_u[pos] + t_u[pos - W]) / P.dX2 - Htdu[pos] * Kx[pos] / (Hm[pos] * (P.dXbp));
            B = P.HaiChiadT + (Htdu[pos] * Kx[pos] + Htdu[pos - W] * Kx[pos - W]) / (Hm[pos] * (P.dXbp));
            if ((Hm[pos - 1] > P.H_TINH) && (Hm[pos + 2] > P.H_TINH)) p = (FS[pos + 1] - FS[pos]);
            if ((Hm[pos - 2] > P.H_TINH) && (Hm[pos + 1] > P.H_TINH)) q = (FS[pos] - FS[pos - 1]);
            p = (1 / (Hm[pos] * P.dYbp)) * (Htdv[pos] * Ky[pos] * p - Htdv[pos - 1] * Ky[pos - 1] * q);
            if ((Hm[pos - 2] > P.H_TINH) && (Hm[pos + 2] > P.H_TINH)) {
                q = (FS[pos + 1] - FS[pos - 1]) / P.dY2;
                q = (t_v[pos] + t_v[pos - 1]) * 0.5 * q * gamav;
            } else q = 0;
            D = FS[pos] / (P.dT * 0.5) - q + p + (S / Hm[pos]);
            if (i == first + 1) {
                if ((br1) || (t_u[first * W + j] == 0)) B += A;
                else D -= A * FS[first * W + j];
            }
            if (i == last - 1) {
                if ((br2) && (t_u[last * W + j] == 0)) B += C;
                else D -= C * FS[last * W + j];
            }
        } else {
            A = -gamav * 0.5 * (t_v[pos - 1] + t_v[pos]) / P.dY2 - Htdv[pos - 1] * Ky[pos - 1] / (Hm[pos] * P.dYbp);
            C = gamav * 0.5 * (t_v[pos - 1] + t_v[pos]) / P.dY2 - Htdv[pos] * Ky[pos] / (Hm[pos] * P.dYbp);
            B = P.HaiChiadT + (Htdv[pos] * Ky[pos] + Htdv[pos - 1] * Ky[pos - 1]) / (Hm[pos] * P.dYbp);
            if ((Hm[pos + 2 * W] > P.H_TINH) && (Hm[pos - W] > P.H_TINH)) p = (FS[pos + W] - FS[pos]);
            if (Hm[pos - 2 * W] > P.H_TINH && Hm[pos + W] > P.H_TINH) q = (FS[pos] - FS[pos - W]);
            p = (1 / (Hm[pos] * P.dXbp)) * (Htdu[pos] * Kx[pos] * p - Htdu[pos - W] * Kx[pos - W] * q);
            if (Hm[pos + 2 * W] > P.H_TINH && Hm[pos - 2 * W] > P.H_TINH) {
                q = (FS[pos + W] - FS[pos - W]) / P.dX2;
                q = (t_u[pos] + t_u[pos - W]) * 0.5 * q * gamav;
            } else q = 0;
            D = FS[pos] / (P.dT * 0.5) - q + p + (S / Hm[pos]);
            if (j == first + 1) {
                if ((br1) || (t_v[i * W + first] == 0)) B = B + A;
                else D = D - A * FS[i * W + first];
            }
            if (j == last - 1) {
                if ((br2) || (t_v[i * W + last] == 0)) B = B + C;
                else D = D - C * FS[i * W + last];
            }
        }
    FsRow w;
    w.A = A; w.B = B; w.C = C; w.D = D;
    return w;
}

/* Concentration rows: _FSj_calculate__mactrix_coeff (Sediment_Transport.cu:205-302,
 * column lines, gamav before Zf and && at the last row, as written) or
 * _FSi_calculate__mactrix_coeff (:62-165, row lines).  Row of cell k sits at
 * line*tw + k. */
template <bool isJ>
__global__ void __launch_bounds__(256) k_fs_coef(const HdDev* __restrict__ dv, double s_start, int ketdinh)
{
    const HdDevG& d = dev_view(dv);
    const double* tdev = d.tdev;
    __shared__ int s_moc[64];
    __shared__ int s_dau[64][24];
    __shared__ int s_cuoi[64][24];
    int N = d.N, M = d.M, SL = d.SL;
    int tx = threadIdx.x, ty = threadIdx.y;
    int j0 = blockIdx.x * 64, i0 = blockIdx.y * 16;
    int tid = ty * 64 + tx;
    int sidx = 3, eidx = isJ ? M - 1 : N - 1;
    if (isJ) {
        for (int t = tid; t < 64; t += 256) {
            int L = j0 + t;
            s_moc[t] = (L >= sidx && L <= eidx) ? min(d.mocj[L], SL) : 0;
        }
        __syncthreads();
        for (int t = tid; t < 64 * SL; t += 256) {
            int c = t / SL, s = t % SL;
            if (s < s_moc[c]) { s_dau[c][s] = d.dauj[(j0 + c) * SL + s]; s_cuoi[c][s] = d.cuoij[(j0 + c) * SL + s]; }
        }
    } else {
        for (int t = tid; t < 16; t += 256) {
            int L = i0 + t;
            s_moc[t] = (L >= sidx && L <= eidx) ? min(d.moci[L], SL) : 0;
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
    const double t = tdev[0];
    const int tw = 2 * (N > M ? N : M) + 4;
    for (int r = ty; r < 16; r += 4) {
        int i = i0 + r;
        if (i > N) break;
        int li = isJ ? tx : r;
        int k = isJ ? i : j;
        int line = isJ ? j : i;
        int first = 0, last = 0, n = s_moc[li];
        bool found = false;
        for (int s = 0; s < n; ++s) {
            if (s_dau[li][s] <= k && k <= s_cuoi[li][s]) { first = s_dau[li][s]; last = s_cuoi[li][s]; found = true; break; }
        }
        if (!found || last < first + 2 || k < first + 1 || k > last - 1) continue;
        bool br1, br2;
        wall_flags2(d, isJ, line, first, last, br1, br2);
        FsRow w = fs_row<isJ>(d, t, s_start, ketdinh, i, j, first, last, br1, br2);
        double A = w.A, B = w.B, C = w.C, D = w.D;
        size_t o = (size_t)line * tw + k;
        d.AA[o] = A; d.BB[o] = B; d.CC[o] = C; d.DD[o] = D;
    }
}

/* tridiagSolver (jump 1) + FSj/FSi_extract_solution (Sediment_Transport.cu:
 * 305-338 / 168-202), one lane per line.  Runs with last - first < 2 have no
 * rows and are skipped. */
template <bool isJ>
__global__ void __launch_bounds__(64) k_fs_solve(const HdDev* __restrict__ dv)
{
    const HdDevG& d = dev_view(dv);
    int N = d.N, M = d.M, W = d.W, SL = d.SL;
    int line = 3 + blockIdx.x * blockDim.x + threadIdx.x;
    if (line > (isJ ? M - 1 : N - 1)) return;
    const int tw = 2 * (N > M ? N : M) + 4;
    int nseg = isJ ? d.mocj[line] : d.moci[line];
    if (nseg > SL) nseg = SL;
    const int* dau = (isJ ? d.dauj : d.daui) + line * SL;
    const int* cuoi = (isJ ? d.cuoij : d.cuoii) + line * SL;
    size_t base = (size_t)line * tw;
    const double *AA = d.AA + base, *BB = d.BB + base, *CC = d.CC + base, *DD = d.DD + base;
    double *Ap = d.Ap + base, *Bp = d.Bp + base;
    auto at = [&](int k) -> size_t { return isJ ? (size_t)k * W + line : (size_t)line * W + k; };
    for (int s = 0; s < nseg; ++s) {
        int first = dau[s], last = cuoi[s];
        if (last < first + 2) continue;
        bool br1, br2;
        wall_flags2(d, isJ, line, first, last, br1, br2);
        int sn = last - first - 2;
        int o = first + 1;
        double x;
        if (sn == 0) x = DD[o] / BB[o];
        else {
            double ap = -CC[o] / BB[o], bp = DD[o] / BB[o];
            Ap[o] = ap; Bp[o] = bp;
            for (int r = 1; r < sn; ++r) {
                double a = AA[o + r];
                double e = a * ap + BB[o + r];
                double nap = -CC[o + r] / e;
                double nbp = (DD[o + r] - (a * bp)) / e;
                ap = nap; bp = nbp;
                Ap[o + r] = ap; Bp[o + r] = bp;
            }
            double a = AA[o + sn];
            x = (DD[o + sn] - (a * bp)) / (BB[o + sn] + (a * ap));
        }
        double vfirst = 0, vlast = 0;
        for (int r = sn; r >= 0; --r) {
            if (r < sn) x = Bp[o + r] + (Ap[o + r] * x);
            double v = x;
            if (v < 0) v = d.P.NDnen;
            d.tFS[at(o + r)] = v;
            if (r == 0) vfirst = v;
            if (r == sn) vlast = v;
        }
        /* end copies (FSj: t_u at first / last; FSi: t_v at first / last) */
        const double* tq = isJ ? d.t_u : d.t_v;
        if ((br1) || (tq[at(first)] == 0)) d.tFS[at(first)] = vfirst;
        else d.tFS[at(first)] = d.FS[at(first)];
        if ((br2) || (tq[at(last)] == 0)) d.tFS[at(last)] = vlast;
        else d.tFS[at(last)] = d.FS[at(last)];
    }
}

/* Fused concentration sweep, one wave per line: rows of 64 interior cells
 * at a time in parallel into LDS, lane 0 runs the Thomas recurrence
 * (UVZSolver_multithread.cu:203-228, jump 1) over them, then back
 * substitution and the extraction into tFS with the end copies
 * (Sediment_Transport.cu:305-338 / 168-202), lanes in parallel. */
template <bool isJ>
__global__ void __launch_bounds__(64) k_fs_line(const HdDev* __restrict__ dv, double s_start, int ketdinh)
{
    const HdDevG& d = dev_view(dv);
    const double* tdev = d.tdev;
    constexpr int CH = 64;
    constexpr int kCap = 2048;
    __shared__ double sA[CH], sB[CH], sC[CH], sD[CH];
    __shared__ double sAp[kCap], sBp[kCap];
    int N = d.N, M = d.M, W = d.W, SL = d.SL;
    int line = 3 + blockIdx.x;
    if (line > (isJ ? M - 1 : N - 1)) return;
    int lane = threadIdx.x;
    const int tw = 2 * (N > M ? N : M) + 4;
    const double t = tdev[0];
    int nseg = isJ ? d.mocj[line] : d.moci[line];
    if (nseg > SL) nseg = SL;
    const int* dau = (isJ ? d.dauj : d.daui) + line * SL;
    const int* cuoi = (isJ ? d.cuoij : d.cuoii) + line * SL;
    double* gAp = d.Ap + (size_t)line * tw;
    double* gBp = d.Bp + (size_t)line * tw;
    auto at = [&](int k) -> size_t { return isJ ? (size_t)k * W + line : (size_t)line * W + k; };
    for (int s = 0; s < nseg; ++s) {
        int first = dau[s], last = cuoi[s];
        if (last < first + 2) continue;
        bool br1, br2;
        wall_flags2(d, isJ, line, first, last, br1, br2);
        int sn = last - first - 2;
        int k0 = first + 1;
        double* oAp = gAp + k0;
        double* oBp = gBp + k0;
        double ap = 0, bp = 0, xs = 0;
        for (int c0 = 0; c0 <= sn; c0 += CH) {
            int r = c0 + lane;
            if (r <= sn) {
                int k = k0 + r;
                int i = isJ ? k : line, j = isJ ? line : k;
                FsRow w = fs_row<isJ>(d, t, s_start, ketdinh, i, j, first, last, br1, br2);
                sA[lane] = w.A; sB[lane] = w.B; sC[lane] = w.C; sD[lane] = w.D;
            }
            __syncthreads();
            if (lane == 0) {
                int re = c0 + CH - 1;
                if (re > sn) re = sn;
                for (int q = c0; q <= re; ++q) {
                    double A = sA[q - c0], B = sB[q - c0], C = sC[q - c0], D = sD[q - c0];
                    if (q == 0 && sn == 0) xs = D / B;
                    else if (q == 0) { ap = -C / B; bp = D / B; }
                    else if (q < sn) {
                        double e = A * ap + B;
                        double nap = -C / e;
                        double nbp = (D - (A * bp)) / e;
                        ap = nap; bp = nbp;
                    } else xs = (D - (A * bp)) / (B + (A * ap));
                    if (q < sn) {
                        if (q < kCap) { sAp[q] = ap; sBp[q] = bp; }
                        else { oAp[q] = ap; oBp[q] = bp; }
                    }
                }
            }
            __syncthreads();
        }
        if (lane == 0) {
            double x = xs;
            if (sn < kCap) sBp[sn] = x; else oBp[sn] = x;
            for (int q = sn - 1; q >= 0; --q) {
                double apq = q < kCap ? sAp[q] : oAp[q];
                double bpq = q < kCap ? sBp[q] : oBp[q];
                x = bpq + (apq * x);
                if (q < kCap) sBp[q] = x; else oBp[q] = x;
            }
        }
        __syncthreads();
        for (int q = lane; q <= sn; q += 64) {
            double v = q < kCap ? sBp[q] : oBp[q];
            if (v < 0) v = d.P.NDnen;
            d.tFS[at(k0 + q)] = v;
        }
        __syncthreads();
        if (lane == 0) {
            double vfirst = d.tFS[at(k0)], vlast = d.tFS[at(last - 1)];
            const double* tqq = isJ ? d.t_u : d.t_v;
            if ((br1) || (tqq[at(first)] == 0)) d.tFS[at(first)] = vfirst;
            else d.tFS[at(first)] = d.FS[at(first)];
            if ((br2) || (tqq[at(last)] == 0)) d.tFS[at(last)] = vlast;
            else d.tFS[at(last)] = d.FS[at(last)];
        }
        __syncthreads();
    }
}

/* The FS sweep with one wave per wet run from the packed run tables (runs_u
 * for the column sweep FSj, runs_v for the row sweep FSi), lines 3 .. M-1 /
 * N-1 as the reference launches them.  Rows are assembled by the 64 lanes in
 * chunks of 256 (four per lane) into LDS; lane 0 carries the Thomas
 * recurrence (the same operations in the same order as k_fs_line and the
 * reference's tridiagSolver); factors of all but the last chunk go to the
 * run's slice of Ap/Bp.  The end copies use the clamped first / last
 * solutions kept in LDS. */
template <bool isJ>
__global__ void __launch_bounds__(64) k_fs_run(const HdDev* __restrict__ dv, double s_start, int ketdinh)
{
    const HdDevG& d = dev_view(dv);
    constexpr int RW = 256;
    __shared__ double sA[RW], sB[RW], sC[RW], sD[RW];
    __shared__ double sV[2]; /* clamped x of the first and last row */
    const int N = d.N, M = d.M, W = d.W;
    const int lane = threadIdx.x;
    const int tw = 2 * (N > M ? N : M) + 4;
    const int lmax = isJ ? M - 1 : N - 1;
    const int cap = (isJ ? M + 3 : N + 3) * d.SL;
    const int cnt = d.run_cnt[isJ ? 0 : 1] < cap ? d.run_cnt[isJ ? 0 : 1] : cap;
    const int4* runs = isJ ? d.runs_u : d.runs_v;
    const double t = d.tdev[0];
    const RunSlice sl = run_slice(cnt);
    for (int ri = sl.start; ri < sl.end; ri += sl.step) {
        const int4 rn = runs[ri];
        const int line = rn.x, first = rn.y, last = rn.z;
        if (line < 3 || line > lmax || last < first + 2) continue;
        auto at = [&](int k) -> size_t { return isJ ? (size_t)k * W + line : (size_t)line * W + k; };
        bool br1, br2;
        wall_flags2(d, isJ, line, first, last, br1, br2);
        const int sn = last - first - 2;
        const int k0 = first + 1;
        double* gAp = d.Ap + (size_t)line * tw + k0;
        double* gBp = d.Bp + (size_t)line * tw + k0;
        const int nch = sn / RW + 1;
        double ap = 0, bp = 0, xs = 0;
        for (int c = 0; c < nch; ++c) {
            const int rb = RW * c;
            const int re = (rb + RW - 1 < sn) ? rb + RW - 1 : sn;
            for (int r = rb + lane; r <= re; r += 64) {
                const int k = k0 + r;
                const int i = isJ ? k : line, j = isJ ? line : k;
                FsRow w = fs_row<isJ>(d, t, s_start, ketdinh, i, j, first, last, br1, br2);
                sA[r - rb] = w.A; sB[r - rb] = w.B; sC[r - rb] = w.C; sD[r - rb] = w.D;
            }
            __syncthreads();
            if (lane == 0) {
                int q = 0;
                if (rb == 0) {
                    if (sn == 0) {
                        xs = sD[0] / sB[0];
                    } else {
                        ap = -sC[0] / sB[0];
                        bp = sD[0] / sB[0];
                        sA[0] = ap; sB[0] = bp;
                    }
                    q = 1;
                }
                const int qmid = ((re < sn - 1) ? re : sn - 1) - rb;
#pragma unroll 4
                for (; q <= qmid; ++q) {
                    double A = sA[q], B = sB[q], C = sC[q], D = sD[q];
                    double e = A * ap + B;
                    double nap = -C / e;
                    double nbp = (D - (A * bp)) / e;
                    ap = nap;
                    bp = nbp;
                    sA[q] = ap; sB[q] = bp;
                }
                if (re == sn && sn > 0) {
                    double A = sA[sn - rb], B = sB[sn - rb], D = sD[sn - rb];
                    xs = (D - (A * bp)) / (B + (A * ap));
                }
            }
            __syncthreads();
            if (c + 1 < nch)
                for (int q = lane; q <= re - rb; q += 64) { gAp[rb + q] = sA[q]; gBp[rb + q] = sB[q]; }
        }
        for (int c = nch - 1; c >= 0; --c) {
            const int rb = RW * c;
            const int re = (rb + RW - 1 < sn) ? rb + RW - 1 : sn;
            if (c + 1 < nch) {
                __syncthreads();
                for (int q = lane; q <= re - rb; q += 64) { sA[q] = gAp[rb + q]; sB[q] = gBp[rb + q]; }
                __syncthreads();
            }
            if (lane == 0) {
                int q = re - rb;
                if (c == nch - 1) { sD[q] = xs; --q; }
#pragma unroll 4
                for (; q >= 0; --q) {
                    xs = sB[q] + (sA[q] * xs);
                    sD[q] = xs;
                }
            }
            __syncthreads();
            for (int q = lane; q <= re - rb; q += 64) {
                double v = sD[q];
                if (v < 0) v = d.P.NDnen;
                d.tFS[at(k0 + rb + q)] = v;
                if (rb + q == 0) sV[0] = v;
                if (rb + q == sn) sV[1] = v;
            }
        }
        __syncthreads();
        if (lane == 0) {
            const double* tqq = isJ ? d.t_u : d.t_v;
            if ((br1) || (tqq[at(first)] == 0)) d.tFS[at(first)] = sV[0];
            else d.tFS[at(first)] = d.FS[at(first)];
            if ((br2) || (tqq[at(last)] == 0)) d.tFS[at(last)] = sV[1];
            else d.tFS[at(last)] = d.FS[at(last)];
        }
        __syncthreads();
    }
}

/* k_fs_run with the speculative parallel Thomas of hd_thomas.h: rows of a
 * run are assembled in chunks of 256 (four per lane; 8 KB of LDS per
 * workgroup, so that about four workgroups per SIMD fit where a batch has
 * many more runs than the chip has waves), the forward sweep and
 * the back substitution run on all 64 lanes with exact verification, so the
 * bits are those of the serial solve (tridiagSolver with jump 1,
 * UVZSolver_multithread.cu:203-262, via the FS rows of Sediment_Transport.cu:
 * 62-338). */
template <bool isJ>
__global__ void __launch_bounds__(64) k_fs_spec(const HdDev* __restrict__ dv, double s_start, int ketdinh)
{
    const HdDevG& d = dev_view(dv);
    constexpr int RW = 256;
    __shared__ double sA[RW], sB[RW], sC[RW], sD[RW]; /* rows; factors replace C / D, x replaces A */
    __shared__ double sV[2]; /* clamped x of the first and last row */
    const int N = d.N, M = d.M, W = d.W;
    const int lane = threadIdx.x;
    const int tw = 2 * (N > M ? N : M) + 4;
    const int lmax = isJ ? M - 1 : N - 1;
    const int cap = (isJ ? M + 3 : N + 3) * d.SL;
    const int cnt = d.run_cnt[isJ ? 0 : 1] < cap ? d.run_cnt[isJ ? 0 : 1] : cap;
    const int4* runs = isJ ? d.runs_u : d.runs_v;
    const double t = d.tdev[0];
    const RunSlice sl = run_slice(cnt);
    for (int ri = sl.start; ri < sl.end; ri += sl.step) {
        const int4 rn = runs[ri];
        const int line = rn.x, first = rn.y, last = rn.z;
        if (line < 3 || line > lmax || last < first + 2) continue;
        auto at = [&](int k) -> size_t { return isJ ? (size_t)k * W + line : (size_t)line * W + k; };
        bool br1, br2;
        wall_flags2(d, isJ, line, first, last, br1, br2);
        const int sn = last - first - 2;
        const int k0 = first + 1;
        double* gAp = d.Ap + (size_t)line * tw + k0;
        double* gBp = d.Bp + (size_t)line * tw + k0;
        const int nch = sn / RW + 1;
        double ap = 0, bp = 0, xs = 0;
        for (int c = 0; c < nch; ++c) {
            const int rb = RW * c;
            const int re = (rb + RW - 1 < sn) ? rb + RW - 1 : sn;
            for (int r = rb + lane; r <= re; r += 64) {
                const int k = k0 + r;
                const int i = isJ ? k : line, j = isJ ? line : k;
                FsRow w = fs_row<isJ>(d, t, s_start, ketdinh, i, j, first, last, br1, br2);
                sA[r - rb] = w.A; sB[r - rb] = w.B; sC[r - rb] = w.C; sD[r - rb] = w.D;
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
            const int g1 = ((re < sn - 1) ? re : sn - 1) - rb;
            hd_spec::spec_forward<false>(sA, sB, sC, sD, g0, g1, ap, bp);
            if (c == 0 && sn > 0) {
                __syncthreads();
                if (lane == 0) { sC[0] = -sC[0] / sB[0]; sD[0] = sD[0] / sB[0]; }
            }
            if (re == sn && sn > 0) {
                double A = sA[sn - rb], B = sB[sn - rb], D = sD[sn - rb];
                xs = (D - (A * bp)) / (B + (A * ap));
            }
            __syncthreads();
            if (c + 1 < nch) {
                for (int q = lane; q <= re - rb; q += 64) { gAp[rb + q] = sC[q]; gBp[rb + q] = sD[q]; }
                __syncthreads();
            }
        }
        double xc = xs;
        for (int c = nch - 1; c >= 0; --c) {
            const int rb = RW * c;
            const int re = (rb + RW - 1 < sn) ? rb + RW - 1 : sn;
            if (c + 1 < nch) {
                __syncthreads();
                for (int q = lane; q <= re - rb; q += 64) { sC[q] = gAp[rb + q]; sD[q] = gBp[rb + q]; }
                __syncthreads();
            } else if (lane == 0) {
                sA[sn - rb] = xs;
            }
            const int hi = ((re < sn - 1) ? re : sn - 1) - rb;
            xc = hd_spec::spec_backward(sC, sD, sA, hi, xc);
            __syncthreads();
            for (int q = lane; q <= re - rb; q += 64) {
                double v = sA[q];
                if (v < 0) v = d.P.NDnen;
                d.tFS[at(k0 + rb + q)] = v;
                if (rb + q == 0) sV[0] = v;
                if (rb + q == sn) sV[1] = v;
            }
        }
        __syncthreads();
        if (lane == 0) {
            const double* tqq = isJ ? d.t_u : d.t_v;
            if ((br1) || (tqq[at(first)] == 0)) d.tFS[at(first)] = sV[0];
            else d.tFS[at(first)] = d.FS[at(first)];
            if ((br2) || (tqq[at(last)] == 0)) d.tFS[at(last)] = sV[1];
            else d.tFS[at(last)] = d.FS[at(last)];
        }
        __syncthreads();
    }
}

/* Update_FS, Sediment_Transport.cu:620-629. */
__global__ void __launch_bounds__(256) k_update_fs(const HdDev* __restrict__ dv, int tiled)
{
    const HdDevG& d = dev_view(dv);
    int j, i;
    if (tiled) {
        if (!tile_cell(d, i, j)) return;
        if (i < 1 || j < 1) return;
    } else {
        j = blockIdx.x * blockDim.x + threadIdx.x + 1;
        i = blockIdx.y + 1;
    }
    if (i > d.N || j > d.M) return;
    d.FS[i * d.W + j] = d.tFS[i * d.W + j];
}

/* Calculate_Qb, Sediment_Transport.cu:341-414 (qb_mode REFERENCE keeps the
 * `!VTH` test as written, FIXED uses VTH != 0; corner copies braced). */
__global__ void __launch_bounds__(256) k_calc_qb(const HdDev* __restrict__ dv, int qb_mode, int tiled)
{
    const HdDevG& d = dev_view(dv);
    int j, i;
    if (tiled) {
        if (!tile_cell(d, i, j)) return;
        if (i < 2 || j < 2) return;
    } else {
        j = blockIdx.x * blockDim.x + threadIdx.x + 2;
        i = blockIdx.y + 2;
    }
    int N = d.N, M = d.M, W = d.W;
    if (i > N || j > M) return;
    const hd_params& P = d.P;
    double* Qbx = d.Qbx;
    double* Qby = d.Qby;
    int pos = i * W + j;
    double Tob, Toee = P.Toe, Tx = 0;
    double vth = d.VTH[pos];
    bool sel = (qb_mode == HD_QB_REFERENCE) ? (!vth) : (vth != 0);
    if (sel && (d.khouot[pos] == 0)) {
        Tob = P.ro * d.Fw[pos] * (vth * vth);
        if (d.H_moi[pos] > P.ghtoe) Toee = P.Toe * (1 + P.hstoe * ((d.H_moi[pos] - P.ghtoe)));
        if (Tob > Toee) Tx = Tob / Toee - 1;
        Qbx[pos] = 0.053 * hd_pow((P.Sx - 1) * P.g, 0.5) * hd_pow(P.dm, 1.5) * hd_pow(Tx, 2.1) * hd_pow(P.Dxr, -0.3)
                   * (d.t_u[pos] + d.t_u[pos - W]) * 0.5 / vth;
        Qby[pos] = 0.053 * hd_pow((P.Sx - 1) * P.g, 0.5) * hd_pow(P.dm, 1.5) * hd_pow(Tx, 2.1) * hd_pow(P.Dxr, -0.3)
                   * (d.t_v[pos] + d.t_v[pos - 1]) * 0.5 / vth;
    } else {
        Qbx[pos] = 0;
        Qby[pos] = 0;
    }
    double qx = Qbx[pos], qy = Qby[pos];
    if (i == 2) { Qbx[pos - W] = qx; Qby[pos - W] = qy; }
    if (i == N) { Qbx[pos + W] = qx; Qby[pos + W] = qy; }
    if (j == 2) {
        Qbx[pos - 1] = qx; Qby[pos - 1] = qy;
        if (i == 2) { Qbx[1 * W + 1] = qx; Qby[1 * W + 1] = qy; }
        if (i == N) { Qbx[(N + 1) * W + 1] = qx; Qby[(N + 1) * W + 1] = qy; }
    }
    if (j == M) {
        Qbx[pos + 1] = qx; Qby[pos + 1] = qy;
        if (i == 2) { Qbx[1 * W + M + 1] = qx; Qby[1 * W + M + 1] = qy; }
        if (i == N) { Qbx[(N + 1) * W + M + 1] = qx; Qby[(N + 1) * W + M + 1] = qy; }
    }
}

/* _bed_load via BedLoad, Sediment_Transport.cu:416-506 / 630-639: rows
 * [sidx, eidx-1], cells of row runs with last - first >= 2 (t_u[pos > 0]
 * reads t_u[1], as written). */
__global__ void __launch_bounds__(256) k_bedload(const HdDev* __restrict__ dv, int ketdinh, int slo, int eoff, int tiled)
{
    const HdDevG& d = dev_view(dv);
    const int sidx = slo, eidx = d.N - eoff;
    int N = d.N, M = d.M, W = d.W, SL = d.SL;
    int i, j;
    int first = 0, last = 0;
    bool found = false;
    if (tiled) { /* run cells lie in the active tiles; the row's table is wave-uniform */
        if (!tile_cell(d, i, j)) return;
        if (i < sidx || i >= eidx || j < 2 || j > M) return;
        const int moc = min(d.moci[i], SL);
        for (int s = 0; s < moc; ++s) {
            const int a = d.daui[i * SL + s], b = d.cuoii[i * SL + s];
            if (a <= j && j <= b) { first = a; last = b; found = true; break; }
        }
    } else {
        __shared__ int s_moc;
        __shared__ int s_dau[24], s_cuoi[24];
        i = sidx + blockIdx.y;
        if (i >= eidx) return;
        if (threadIdx.x == 0) s_moc = min(d.moci[i], SL);
        __syncthreads();
        if (threadIdx.x < s_moc) { s_dau[threadIdx.x] = d.daui[i * SL + threadIdx.x]; s_cuoi[threadIdx.x] = d.cuoii[i * SL + threadIdx.x]; }
        __syncthreads();
        j = blockIdx.x * blockDim.x + threadIdx.x + 2;
        if (j > M) return;
        for (int s = 0; s < s_moc; ++s)
            if (s_dau[s] <= j && j <= s_cuoi[s]) { first = s_dau[s]; last = s_cuoi[s]; found = true; break; }
    }
    if (!found || last - first < 2) return;
    const hd_params& P = d.P;
    const double *FS = d.FS, *Hm = d.H_moi, *t_u = d.t_u, *t_v = d.t_v, *VTH = d.VTH, *Htdu = d.Htdu, *Htdv = d.Htdv,
                 *Fw = d.Fw, *Kx = d.Kx, *Ky = d.Ky, *Qbx = d.Qbx, *Qby = d.Qby;
    const int* kh = d.khouot;
    int pos = i * W + j;
    double p = 0, q = 0;
    if (kh[pos - W] == 1) {
        if (t_u[pos] < 0) p = (-3 * Qbx[pos] + 4 * Qbx[pos + W] - Qbx[pos + 2 * W]) / P.dX2;
    } else {
        if (kh[pos + W] == 1) {
            if (t_u[1]) p = (3 * Qbx[pos] - 4 * Qbx[pos - W] + Qbx[pos - 2 * W]) / P.dX2;
        } else p = (Qbx[pos + W] - Qbx[pos - W]) / P.dX2;
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
    double Uf = sqrt(P.g) * fabs(VTH[pos]) / c;
    double wsm = P.wss * hd_pow((1 - FS[pos]), 4);
    double Zf = 0, S = 0;
    if (FS[pos] > 0.0) Zf = wsm / (0.4 * (Uf + 2 * wsm));
    if (ketdinh) S = source_chs(P, wsm, Zf, FS[pos], Fw[pos], VTH[pos], Hm[pos]);
    else S = source_nchs(P, wsm, Zf, Uf, Hm[pos], FS[pos]);
    d.dH[pos] += P.dT / (1 - P.Dorong) * (S + tH);
}

} // namespace

static inline dim3 tile_grid(const HdBatch& b, int ntiles) { return dim3((ntiles + 3) / 4, 1, b.n); }

void hd_launch_vth_hesok(const HdBatch& b, int ntiles, hipStream_t s)
{
    if (ntiles > 0) {
        hipLaunchKernelGGL(k_vth_hesok, tile_grid(b, ntiles), dim3(256), 0, s, b.dv, 1);
        return;
    }
    dim3 g2((b.M + 1 + 255) / 256, b.N + 1, b.n);
    hipLaunchKernelGGL(k_vth_hesok, g2, dim3(256), 0, s, b.dv, 0);
}

void hd_launch_fs_coef(const HdBatch& b, int which, double s_start, int ketdinh, hipStream_t s)
{
    dim3 blk(64, 4);
    dim3 grd((b.M + 3 + 63) / 64, (b.N + 3 + 15) / 16, b.n);
    if (which == 0) hipLaunchKernelGGL(k_fs_coef<true>, grd, blk, 0, s, b.dv, s_start, ketdinh);
    else hipLaunchKernelGGL(k_fs_coef<false>, grd, blk, 0, s, b.dv, s_start, ketdinh);
}

void hd_launch_fs_solve(const HdBatch& b, int which, hipStream_t s)
{
    int nl = (which == 0 ? b.M : b.N) - 3;
    if (nl <= 0) return;
    if (which == 0) hipLaunchKernelGGL(k_fs_solve<true>, dim3((nl + 63) / 64, 1, b.n), dim3(64), 0, s, b.dv);
    else hipLaunchKernelGGL(k_fs_solve<false>, dim3((nl + 63) / 64, 1, b.n), dim3(64), 0, s, b.dv);
}

void hd_launch_update_fs(const HdBatch& b, int ntiles, hipStream_t s)
{
    if (ntiles > 0) {
        hipLaunchKernelGGL(k_update_fs, tile_grid(b, ntiles), dim3(256), 0, s, b.dv, 1);
        return;
    }
    hipLaunchKernelGGL(k_update_fs, dim3((b.M + 255) / 256, b.N, b.n), dim3(256), 0, s, b.dv, 0);
}

void hd_launch_calc_qb(const HdBatch& b, int qb_mode, int ntiles, hipStream_t s)
{
    if (ntiles > 0) {
        hipLaunchKernelGGL(k_calc_qb, tile_grid(b, ntiles), dim3(256), 0, s, b.dv, qb_mode, 1);
        return;
    }
    hipLaunchKernelGGL(k_calc_qb, dim3((b.M - 1 + 255) / 256, b.N - 1, b.n), dim3(256), 0, s, b.dv, qb_mode, 0);
}

void hd_launch_bedload(const HdBatch& b, int ketdinh, int slo, int eoff, int ntiles, hipStream_t s)
{
    if (ntiles > 0) {
        hipLaunchKernelGGL(k_bedload, tile_grid(b, ntiles), dim3(256), 0, s, b.dv, ketdinh, slo, eoff, 1);
        return;
    }
    int nr = b.N - eoff - slo;
    if (nr > 0)
        hipLaunchKernelGGL(k_bedload, dim3((b.M - 1 + 255) / 256, nr, b.n), dim3(256), 0, s, b.dv, ketdinh, slo, eoff, 0);
}

void hd_launch_fs_line(const HdBatch& b, int which, double s_start, int ketdinh, hipStream_t s)
{
    int nl = (which == 0 ? b.M : b.N) - 3;
    if (nl <= 0) return;
    if (which == 0) hipLaunchKernelGGL(k_fs_line<true>, dim3(nl, 1, b.n), dim3(64), 0, s, b.dv, s_start, ketdinh);
    else hipLaunchKernelGGL(k_fs_line<false>, dim3(nl, 1, b.n), dim3(64), 0, s, b.dv, s_start, ketdinh);
}

void hd_launch_fs_run(const HdBatch& b, int which, double s_start, int ketdinh, int nblocks, int spec, hipStream_t s)
{
    dim3 g(nblocks, 1, b.n);
    if (spec) {
        if (which == 0) hipLaunchKernelGGL(k_fs_spec<true>, g, dim3(64), 0, s, b.dv, s_start, ketdinh);
        else hipLaunchKernelGGL(k_fs_spec<false>, g, dim3(64), 0, s, b.dv, s_start, ketdinh);
    } else {
        if (which == 0) hipLaunchKernelGGL(k_fs_run<true>, g, dim3(64), 0, s, b.dv, s_start, ketdinh);
        else hipLaunchKernelGGL(k_fs_run<false>, g, dim3(64), 0, s, b.dv, s_start, ketdinh);
    }
}
