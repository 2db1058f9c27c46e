/*
 * hd_api.hip -- the extern "C" boundary of libhydist.so (include/hydist.h):
 * context and device memory, named field transfer, the step sequence of
 * Engine.hydraulic_Calculation (Engine.py:144-366) and its HIP-graph replay.
 */
#include "hd_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace {

/* set the device time of every context of a batch (grid.z = batch) */
__global__ void k_set_t(const HdDev* dv, double t)
{
    if (threadIdx.x == 0) dv[blockIdx.z].tdev[0] = t;
}

struct Field {
    double** dptr;   /* points into hd_ctx::d so buffer swaps are seen */
    int** iptr;
    size_t count;
};

/* Argument_Pointers, constant.cuh:6-20, field for field. */
struct ArgMirror {
    int M, N;
    double hmax_u, hmax_d, hmax_l, hmax_r;
    int *bienQ, *daui, *dauj, *cuoii, *cuoij, *moci, *mocj, *khouot, *boundary_type;
    double *h, *v, *u, *z, *t_u, *t_v, *t_z, *Htdu, *Htdv, *H_moi, *htaiz, *htaiz_bd;
    double *ubt, *ubp, *vbt, *vbd;
    double *hsnham, *VISCOIDX, *Kx1, *Ky1, *Tsyw, *Tsxw;
    double *bc_up, *bc_down, *bc_left, *bc_right;
    double* hi;
    double *FS, *tFS, *CC_u, *CC_d, *CC_l, *CC_r;
    double *VTH, *Kx, *Ky, *Fw;
    double *Qbx, *Qby;
    double* dH;
};

/* Array_Pointers, constant.cuh:22-29.  This library keeps no a1..f5 / x / ep
 * arrays (coefficients stay in registers); those members are null. */
struct ArrMirror {
    double *a1, *b1, *c1, *d1, *a2, *c2, *d2;
    double *f1, *f2, *f3, *f5;
    double *AA, *BB, *CC, *DD;
    double* x;
    double *Ap, *Bp, *ep;
    int* SN;
};

struct GraphKey {
    int sed0, sed1, bed, count, kd, kn, analytic, ketdinh, qb_mode;
    double s_start;
    int ntiles;       /* per-row kernels over this many active tiles (0: full grid) */
    int sed_ntiles;   /* the same for the sediment sweep's pointwise kernels */
    int bed_ntiles;   /* and for the bed-load flux */
    bool operator<(const GraphKey& o) const
    {
        return std::tie(sed0, sed1, bed, count, kd, kn, analytic, ketdinh, qb_mode, s_start, ntiles, sed_ntiles,
                        bed_ntiles) <
               std::tie(o.sed0, o.sed1, o.bed, o.count, o.kd, o.kn, o.analytic, o.ketdinh, o.qb_mode, o.s_start,
                        o.ntiles, o.sed_ntiles, o.bed_ntiles);
    }
};

const char* kNames[] = {"h", "hsnham", "VISCOIDX", "u", "v", "z", "t_u", "t_v", "t_z", "Htdu", "Htdv", "H_moi",
                        "htaiz", "htaiz_bd", "Kx1", "Ky1", "Tsxw", "Tsyw", "FS", "tFS", "VTH", "Kx", "Ky", "Fw",
                        "Qbx", "Qby", "dH"};

} // namespace

/* per-kernel HIP-event timing (hd_profile / hd_kernel_times) */
struct Prof {
    bool on = false;
    struct Rec { const char* name; hipEvent_t a, b; };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    std::vector<std::string> order;
    std::map<std::string, std::pair<double, int>> acc;
};

struct hd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int N = 0, M = 0, SL = 0, series_len = 0, total_time = 0;
    hd_params P;
    HdDev d;                    /* host copy; t_u/tmp_u and t_v/tmp_v swap with the smoothing passes */
    double *t_u0 = nullptr, *t_v0 = nullptr; /* the buffers named t_u / t_v at creation */
    HdDev* dv = nullptr;        /* device copies of d in the four swap states (bit 0: u, bit 1: v) */
    std::map<std::string, Field> fields;
    std::vector<void*> allocs;
    std::vector<std::pair<void*, size_t>> state_allocs;
    void* arg_mirror = nullptr;
    void* arr_mirror = nullptr;
    bool initialised = false;
    std::string err;
    std::map<GraphKey, hipGraphExec_t> graphs;
    std::map<GraphKey, hipGraphExec_t> sed_graphs;   /* sediment-only steps (hd_sediment_run) */
    /* hydrodynamic solver form (HYDIST_SOLVER): 0 = one wave per wet run,
     * speculative parallel Thomas (default), 3 = one wave per wet run, serial
     * Thomas on one lane ("run"), 1 = one wave per line ("line"), 2 = rows
     * kernel + one lane per line ("rows").  All four give identical bits. */
    int solver = 0;
    bool solver_v1 = false;     /* two-kernel sediment form as well (HYDIST_SOLVER=rows) */
    /* wetting form (HYDIST_WETTING): 0 default, 1 candidate windows for every
     * line ("window"), 2 whole-line scans ("scan"); identical bits */
    int wet_mode = 0;
    int run_blocks_u = 0, run_blocks_v = 0;
    double* bed_snap = nullptr; /* dH snapshot for hd_bed_extrapolate (allocated on first use) */
    double* z_snap = nullptr;   /* z snapshot for hd_z_change (allocated on first use), + 5 result slots */
    int ntiles = 0;             /* active row tiles (computed at init / prepare) */
    bool settled = false;       /* a full step has run since init / upload: land tiles are at a fixed point */
    bool sed_settled = false;   /* ... and a sediment sweep over the whole grid (VTH, hesoK, FS = tFS) */
    bool bed_settled = false;   /* ... and a bed-load flux over the whole grid */
    int rb_mult = 2;            /* solver blocks per line (autotuned) */
    int bw = 256;               /* block width of the per-row full-grid kernels (autotuned) */
    Prof prof;
    int state() const { return (d.t_u != t_u0 ? 1 : 0) | (d.t_v != t_v0 ? 2 : 0); }
};

/* Where a step is enqueued: one context (n = 1) or an ensemble of contexts
 * on one device.  dv holds n device HdDev copies per swap state, state-major;
 * a launch passes the copies of the current state. */
struct Exec {
    hipStream_t stream;
    const HdDev* dv;
    int n, N, M;
    int state;
    int solver;
    bool solver_v1;
    int wet_mode;
    int run_blocks_u, run_blocks_v;
    int bw;
    Prof* prof;
    hd_ctx* single;   /* mirrors the swaps into the context's host copy */
    HdBatch b() const { return HdBatch{dv + (size_t)state * n, n, N, M}; }
    void swap_u()
    {
        state ^= 1;
        if (single) std::swap(single->d.t_u, single->d.tmp_u);
    }
    void swap_v()
    {
        state ^= 2;
        if (single) std::swap(single->d.t_v, single->d.tmp_v);
    }
};

/* solver grid: rb_mult blocks per line, at most one per table entry, in
 * whole rounds of the eight XCDs (the solvers deal table slices by XCD) */
static void set_knobs(hd_ctx* c, int rb_mult, int bw)
{
    int N = c->N, M = c->M, SL = c->SL;
    c->rb_mult = rb_mult;
    c->bw = bw;
    c->run_blocks_u = rb_mult * (M + 3) < (M + 3) * SL ? rb_mult * (M + 3) : (M + 3) * SL;
    c->run_blocks_v = rb_mult * (N + 3) < (N + 3) * SL ? rb_mult * (N + 3) : (N + 3) * SL;
    c->run_blocks_u = (c->run_blocks_u + 7) & ~7;
    c->run_blocks_v = (c->run_blocks_v + 7) & ~7;
}

static Exec exec_of(hd_ctx* c)
{
    return Exec{c->stream, c->dv, 1, c->N, c->M, c->state(), c->solver, c->solver_v1, c->wet_mode, c->run_blocks_u,
                c->run_blocks_v, c->bw, &c->prof, c};
}

static hipEvent_t take_event(Prof* p)
{
    if (!p->pool.empty()) {
        hipEvent_t e = p->pool.back();
        p->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

/* Run one launcher, bracketed by events on the stream when profiling is on. */
template <class F>
static void launch(Exec& x, const char* name, F&& f)
{
    if (!x.prof || !x.prof->on) { f(); return; }
    Prof::Rec r{name, take_event(x.prof), take_event(x.prof)};
    (void)hipEventRecord(r.a, x.stream);
    f();
    (void)hipEventRecord(r.b, x.stream);
    x.prof->pending.push_back(r);
}

static thread_local std::string g_err;

static int fail(hd_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    g_err = msg;
    return code;
}

#define HIPCHK(c, x)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) return fail((c), HD_EHIP, std::string(#x ": ") + hipGetErrorString(e_)); \
    } while (0)

/* Windx/Windy/Ufr/Dxr/wss, constant.cuh:126-133, evaluated once on the host
 * with the project's math library. */
static void derive_params(hd_params& p)
{
    double aw = std::fabs(p.Wind);
    p.Windx = 0.0013 * (0.00075 + 0.000067 * aw) * aw * p.Wind * hd_cos(p.huonggio * (p.PI / 180));
    p.Windy = 0.0013 * (0.00075 + 0.000067 * aw) * aw * p.Wind * hd_sin(p.huonggio * (p.PI / 180));
    p.Ufr = 0.25 * hd_pow((p.Sx - 1) * p.g, 8.0 / 15.0) * hd_pow(p.dm, 9.0 / 15.0) * hd_pow(p.muy, -1.0 / 15.0);
    p.Dxr = p.dm * hd_pow(p.g * (p.Sx - 1) / p.muy * p.muy, (1.0 / 3.0));
    p.wss = 10 * p.muy / p.dm * (std::sqrt(1 + 0.01 * (p.Sx - 1) * 9.81 * hd_pow(p.dm, 3) / hd_pow(p.muy, 2)) - 1);
}

/* device allocation of a context, zero-filled; `state` allocations hold
 * data (copied by hd_clone), the others hold pointer tables.  The fill runs
 * on the context's own stream and has finished when this returns: the
 * stream is non-blocking, so a fill on the null stream would not be ordered
 * before the copies and kernels queued on it next (a clone's state copy, a
 * snapshot) and could land after them. */
static int dalloc(hd_ctx* c, void** p, size_t bytes, bool state = true)
{
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(c, HD_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->allocs.push_back(*p);
    if (state) c->state_allocs.push_back(std::make_pair(*p, bytes));
    e = hipMemsetAsync(*p, 0, bytes, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, HD_EHIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    return HD_OK;
}

/* Device HdDev of a context in its current swap state, its stream and the
 * state (used by the surrogate's packing kernels in hd_cnn.hip). */
const HdDev* hd_ctx_device_view(hd_ctx* c, hipStream_t* stream, int* state)
{
    if (stream) *stream = c->stream;
    if (state) *state = c->state();
    return c->dv + c->state();
}

extern "C" {

const char* hd_version(void) { return "hydist-mi355x 0.1 (gfx950, FP64 ADI)"; }

const char* hd_error_string(int code)
{
    switch (code) {
    case HD_OK: return "ok";
    case HD_EINVAL: return "invalid argument";
    case HD_ESHAPE: return "byte count does not match the field layout";
    case HD_ENAME: return "unknown field name";
    case HD_EHIP: return "HIP runtime error";
    case HD_ESEGOVERFLOW: return "a line has more wet runs than segment_limit";
    case HD_ESTATE: return "call out of order";
    case HD_ERANGE: return "boundary series too short for the time";
    case HD_ENOMEM: return "out of device memory";
    default: return "unknown error";
    }
}

const char* hd_last_error(const hd_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

int hd_create(int device, const hd_params* p, int N, int M, int series_len, int total_time, hd_ctx** out)
{
    if (!p || !out || N < 4 || M < 4 || p->segment_limit < 1 || p->segment_limit > 24 || series_len < 2 ||
        total_time < 2)
        return fail(nullptr, HD_EINVAL, "hd_create: bad arguments (N, M >= 4; 1 <= segment_limit <= 24)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(nullptr, HD_EINVAL, "hd_create: no such HIP device");
    hd_ctx* c = new hd_ctx();
    c->device = device;
    c->N = N; c->M = M; c->SL = p->segment_limit; c->series_len = series_len; c->total_time = total_time;
    c->P = *p;
    derive_params(c->P);
    {
        const char* sv = getenv("HYDIST_SOLVER");
        c->solver_v1 = sv && std::string(sv) == "rows";
        c->solver = c->solver_v1 ? 2 : (sv && std::string(sv) == "line") ? 1 : (sv && std::string(sv) == "run") ? 3 : 0;
        const char* wv = getenv("HYDIST_WETTING");
        c->wet_mode = (wv && std::string(wv) == "window") ? 1 : (wv && std::string(wv) == "scan") ? 2 : 0;
    }
    int rc;
    if (hipSetDevice(device) != hipSuccess) { delete c; return fail(nullptr, HD_EHIP, "hipSetDevice failed"); }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(nullptr, HD_EHIP, "hipStreamCreate failed");
    }
    HdDev& d = c->d;
    std::memset(&d, 0, sizeof(d));
    d.N = N; d.M = M; d.W = M + 3; d.SL = c->SL; d.total_time = total_time; d.series_len = series_len;
    d.P = c->P;
    size_t F = (size_t)(N + 3) * (M + 3);
    double** f2d[] = {&d.h, &d.hsnham, &d.VISCOIDX, &d.u, &d.v, &d.z, &d.t_u, &d.t_v, &d.t_z, &d.Htdu, &d.Htdv,
                      &d.H_moi, &d.htaiz, &d.htaiz_bd, &d.Kx1, &d.Ky1, &d.Tsxw, &d.Tsyw, &d.FS, &d.tFS, &d.VTH,
                      &d.Kx, &d.Ky, &d.Fw, &d.Qbx, &d.Qby, &d.dH};
    for (size_t k = 0; k < sizeof(f2d) / sizeof(f2d[0]); ++k) {
        if ((rc = dalloc(c, (void**)f2d[k], F * 8))) { hd_destroy(c); return rc; }
        c->fields[kNames[k]] = Field{f2d[k], nullptr, F};
    }
    c->fields["VISCOINDX"] = c->fields["VISCOIDX"];
    if ((rc = dalloc(c, (void**)&d.khouot, F * 4))) { hd_destroy(c); return rc; }
    c->fields["khouot"] = Field{nullptr, &d.khouot, F};
    struct OneD { const char* n; double** p; size_t cnt; } od[] = {
        {"ubt", &d.ubt, (size_t)M + 2}, {"ubp", &d.ubp, (size_t)M + 2}, {"vbt", &d.vbt, (size_t)N + 2},
        {"vbd", &d.vbd, (size_t)N + 2}, {"hi", &d.hi, (size_t)2 * (M + N + 6)},
        {"bc_up", &d.bc_up, (size_t)series_len}, {"bc_down", &d.bc_down, (size_t)series_len},
        {"bc_left", &d.bc_left, (size_t)series_len}, {"bc_right", &d.bc_right, (size_t)series_len},
        {"CC_u", &d.CC_u, (size_t)series_len}, {"CC_d", &d.CC_d, (size_t)series_len},
        {"CC_l", &d.CC_l, (size_t)series_len}, {"CC_r", &d.CC_r, (size_t)series_len}};
    for (auto& o : od) {
        if ((rc = dalloc(c, (void**)o.p, o.cnt * 8))) { hd_destroy(c); return rc; }
        c->fields[o.n] = Field{o.p, nullptr, o.cnt};
    }
    int SL = c->SL;
    struct OneI { const char* n; int** p; size_t cnt; } oi[] = {
        {"bienQ", &d.bienQ, 4}, {"boundary_type", &d.boundary_type, (size_t)4 * SL},
        {"daui", &d.daui, (size_t)(N + 2) * SL}, {"cuoii", &d.cuoii, (size_t)(N + 2) * SL},
        {"dauj", &d.dauj, (size_t)(M + 2) * SL}, {"cuoij", &d.cuoij, (size_t)(M + 2) * SL},
        {"moci", &d.moci, (size_t)N + 2}, {"mocj", &d.mocj, (size_t)M + 2}};
    for (auto& o : oi) {
        if ((rc = dalloc(c, (void**)o.p, o.cnt * 4))) { hd_destroy(c); return rc; }
        c->fields[o.n] = Field{nullptr, o.p, o.cnt};
    }
    int L = (N > M ? N : M);
    size_t tw = 2 * (size_t)L + 4, nlines = (size_t)L + 3;
    if ((rc = dalloc(c, (void**)&d.hmax, 4 * 8))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&d.tmp_u, F * 8))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&d.tmp_v, F * 8))) { hd_destroy(c); return rc; }
    double** rows[] = {&d.AA, &d.BB, &d.CC, &d.DD, &d.Ap, &d.Bp};
    for (auto r : rows)
        if ((rc = dalloc(c, (void**)r, tw * nlines * 8))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&d.seg_end, nlines * SL * 6 * 8))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&d.seg_sn, nlines * SL * 4))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&d.runs_u, (size_t)(M + 3) * SL * sizeof(int4)))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&d.runs_v, (size_t)(N + 3) * SL * sizeof(int4)))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&d.run_cnt, 4 * 4))) { hd_destroy(c); return rc; }
    /* one block per run while the table holds up to twice as many runs as
       lines; blocks stride over the rest */
    set_knobs(c, c->rb_mult, c->bw);
    if ((rc = dalloc(c, (void**)&d.flags, 16 * 4))) { hd_destroy(c); return rc; }
    /* [0]: per-step accumulator of hd_run_counted, [1]: one-shot count of hd_wet_count */
    if ((rc = dalloc(c, (void**)&d.wet_count, 16))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&d.tdev, 8))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&d.cand_cnt, nlines * 4))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&d.cand, nlines * 64 * 4))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&d.dryflag, F))) { hd_destroy(c); return rc; }
    {
        size_t ncb = (size_t)(M + 3 + 63) / 64;
        if ((rc = dalloc(c, (void**)&d.tiles, (size_t)(N + 3) * ncb * 4))) { hd_destroy(c); return rc; }
        if ((rc = dalloc(c, (void**)&d.ntiles, 4))) { hd_destroy(c); return rc; }
        if ((rc = dalloc(c, (void**)&d.wetb, (size_t)(N + 3) * ncb * 8))) { hd_destroy(c); return rc; }
        if ((rc = dalloc(c, (void**)&d.kh0b, (size_t)(N + 3) * ncb * 8))) { hd_destroy(c); return rc; }
    }
    c->t_u0 = d.t_u;
    c->t_v0 = d.t_v;
    if ((rc = dalloc(c, (void**)&c->dv, 4 * sizeof(HdDev), false))) { hd_destroy(c); return rc; }
    {
        HdDev st[4];
        for (int k = 0; k < 4; ++k) {
            st[k] = d;
            if (k & 1) std::swap(st[k].t_u, st[k].tmp_u);
            if (k & 2) std::swap(st[k].t_v, st[k].tmp_v);
        }
        if (hipMemcpy(c->dv, st, sizeof(st), hipMemcpyHostToDevice) != hipSuccess) {
            hd_destroy(c);
            return fail(nullptr, HD_EHIP, "hipMemcpy device state copies");
        }
    }
    if ((rc = dalloc(c, &c->arg_mirror, sizeof(ArgMirror), false))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, &c->arr_mirror, sizeof(ArrMirror), false))) { hd_destroy(c); return rc; }
    /* the reference's host arrays start as zeros except khouot = 2
       (Global_Variables.py:93-127) */
    std::vector<int> two(F, 2);
    if (hipMemcpy(d.khouot, two.data(), F * 4, hipMemcpyHostToDevice) != hipSuccess) {
        hd_destroy(c);
        return fail(nullptr, HD_EHIP, "hipMemcpy khouot");
    }
    *out = c;
    return HD_OK;
}

int hd_destroy(hd_ctx* c)
{
    if (!c) return HD_OK;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto& g : c->graphs) hipGraphExecDestroy(g.second);
    for (auto& g : c->sed_graphs) hipGraphExecDestroy(g.second);
    for (auto& r : c->prof.pending) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
    for (auto e : c->prof.pool) hipEventDestroy(e);
    for (void* p : c->allocs) hipFree(p);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return HD_OK;
}

int hd_get_params(const hd_ctx* c, hd_params* out)
{
    if (!c || !out) return HD_EINVAL;
    *out = c->P;
    return HD_OK;
}

static int lookup(hd_ctx* c, const char* name, void** ptr, size_t* nbytes, int* dtype)
{
    if (!c || !name) return fail(c, HD_EINVAL, "null argument");
    auto it = c->fields.find(name);
    if (it == c->fields.end()) return fail(c, HD_ENAME, std::string("unknown field ") + name);
    const Field& f = it->second;
    if (f.dptr) { *ptr = *f.dptr; *nbytes = f.count * 8; *dtype = 0; }
    else { *ptr = *f.iptr; *nbytes = f.count * 4; *dtype = 1; }
    return HD_OK;
}

int hd_field_info(hd_ctx* c, const char* name, size_t* nbytes, int* dtype)
{
    void* p;
    size_t nb;
    int dt;
    int rc = lookup(c, name, &p, &nb, &dt);
    if (rc) return rc;
    if (nbytes) *nbytes = nb;
    if (dtype) *dtype = dt;
    return HD_OK;
}

static int check_flags(hd_ctx* c)
{
    int f[2];
    HIPCHK(c, hipMemcpyAsync(f, c->d.flags, sizeof(f), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (f[0]) return fail(c, HD_ESEGOVERFLOW, "a line has more wet runs than segment_limit");
    if (f[1]) return fail(c, HD_ERANGE, "boundary series too short for the time");
    return HD_OK;
}

static int compute_tiles(hd_ctx* c);

int hd_upload(hd_ctx* c, const char* name, const void* src, size_t nbytes)
{
    void* p;
    size_t nb;
    int dt;
    int rc = lookup(c, name, &p, &nb, &dt);
    if (rc) return rc;
    if (!src || nbytes != nb) return fail(c, HD_ESHAPE, std::string("size mismatch for ") + name);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(p, src, nb, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->settled = false;
    c->sed_settled = false;
    c->bed_settled = false;
    /* a new mask can open cells outside the active tiles */
    if (c->initialised && std::strcmp(name, "khouot") == 0) return compute_tiles(c);
    return HD_OK;
}

int hd_download(hd_ctx* c, const char* name, void* dst, size_t nbytes)
{
    void* p;
    size_t nb;
    int dt;
    int rc = lookup(c, name, &p, &nb, &dt);
    if (rc) return rc;
    if (!dst || nbytes != nb) return fail(c, HD_ESHAPE, std::string("size mismatch for ") + name);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(dst, p, nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return check_flags(c);
}

int hd_get_hmax(hd_ctx* c, double out[4])
{
    if (!c || !out) return HD_EINVAL;
    HIPCHK(c, hipMemcpyAsync(out, c->d.hmax, 32, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return HD_OK;
}

/* Active 64-column row tiles: those holding or bordering (one cell) a cell
 * with khouot != 2.  khouot == 2 never changes after Onetime_init (wetting
 * opens only khouot == 1 cells, drying closes only khouot == 0 cells), and
 * outside these tiles no kernel writes a value that differs from the one
 * already there once a full step has run. */
static int compute_tiles(hd_ctx* c)
{
    int N = c->N, M = c->M, W = M + 3, ncb = (M + 3 + 63) / 64;
    std::vector<int> kh((size_t)(N + 3) * W);
    HIPCHK(c, hipMemcpyAsync(kh.data(), c->d.khouot, kh.size() * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<char> live((size_t)(N + 3) * ncb, 0);
    for (int i = 0; i < N + 3; ++i)
        for (int j = 0; j < M + 3; ++j)
            if (kh[(size_t)i * W + j] != 2)
                for (int di = -1; di <= 1; ++di) {
                    int ii = i + di;
                    if (ii < 0 || ii > N + 2) continue;
                    for (int dj = -1; dj <= 1; ++dj) {
                        int jj = j + dj;
                        if (jj < 0 || jj > M + 2) continue;
                        live[(size_t)ii * ncb + jj / 64] = 1;
                    }
                }
    std::vector<int> list;
    for (size_t t = 0; t < live.size(); ++t)
        if (live[t]) list.push_back((int)t);
    int n = (int)list.size();
    if (n) HIPCHK(c, hipMemcpyAsync(c->d.tiles, list.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d.ntiles, &n, 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->ntiles = n;
    c->settled = false;
    c->sed_settled = false;
    c->bed_settled = false;
    return HD_OK;
}

/* rep.py:126-139: Onetime_init, Find_Calculation_limits_x/y, Htuongdoi,
 * preprocess_data. */
int hd_init(hd_ctx* c)
{
    if (!c) return HD_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    HdBatch b = exec_of(c).b();
    hd_launch_onetime_init(b, c->stream);
    hd_launch_find_limits(b, c->stream);
    hd_launch_htuongdoi(b, c->stream);
    hd_launch_preprocess(b, c->stream);
    HIPCHK(c, hipGetLastError());
    int rc = check_flags(c);
    if (rc) return rc;
    if ((rc = compute_tiles(c))) return rc;
    c->initialised = true;
    return HD_OK;
}

/* rep.py:135-139 on the current state (segment finders, Htuongdoi,
 * preprocess_data), for a restart whose u, v, z, khouot, FS were uploaded
 * after hd_init (rep.py:133-134). */
int hd_prepare(hd_ctx* c)
{
    if (!c) return HD_EINVAL;
    if (!c->initialised) return fail(c, HD_ESTATE, "hd_prepare before hd_init");
    HIPCHK(c, hipSetDevice(c->device));
    HdBatch b = exec_of(c).b();
    hd_launch_find_limits(b, c->stream);
    hd_launch_htuongdoi(b, c->stream);
    hd_launch_preprocess(b, c->stream);
    HIPCHK(c, hipGetLastError());
    int rc = check_flags(c);
    if (rc) return rc;
    return compute_tiles(c);
}

static bool bed_trigger(double t, const hd_step_opts& o)
{
    if (o.bed_mode == HD_BED_EVERY_STEP) return t >= o.bed_change_start;
    double it = (double)(long long)t;
    if (t - it != 0) return false;
    if (o.bed_change_start == 0) return false;
    return std::fmod(it, o.bed_change_start) == 0;
}

static bool analytic_mode(const hd_step_opts& o)
{
    bool channel = o.kenhhepd || o.kenhhepng;
    return o.boundary_mode == HD_BND_ANALYTIC || (o.boundary_mode == HD_BND_AUTO && channel);
}

static int check_series(hd_ctx* c, double t, const hd_step_opts& o)
{
    if (analytic_mode(o)) return HD_OK;
    int t1 = (int)(t / 3600);
    if (t1 < 0 || t1 + 1 >= c->total_time || t1 + 1 >= c->series_len)
        return fail(c, HD_ERANGE, "boundary series too short for t");
    return HD_OK;
}

/* One half-step of Engine.py:146-239 (which = 0, UZ) or :260-361 (which = 1,
 * VZ), enqueued on the stream for every context of the batch.  The boundary
 * values of the half-step (Update_Boundary_Value) and the device time step:
 * in the default form with the boundary series (not the analytic channel
 * tide, not the narrow-channel index shifts), the sweep's runs that own the
 * edge cells write the values and the explicit pass after it advances the
 * time; otherwise a k_boundary launch does both.  Returns in *slo / *eoff
 * the row range of a following bed change (rows slo .. N - eoff - 1). */
static void enqueue_half(Exec& x, int which, const hd_step_opts& o, bool sediment, int count_wet, int ntiles,
                         int sed_ntiles, int* slo_out, int* eoff_out)
{
    hipStream_t s = x.stream;
    bool channel = o.kenhhepd || o.kenhhepng;
    int an = analytic_mode(o) ? 1 : 0;
    const int own = (x.solver == 0 && !an && !channel) ? 1 : 0;
    if (!own) launch(x, "boundary", [&] { hd_launch_boundary_t(x.b(), an, s); });
    int slo = 2, eoff = 0;
    if (which == 0) {
        int shift = (channel && o.kenhhepd == 1) ? 1 : 0;
        if (x.solver == 2) {
            launch(x, "adi_coef_uz", [&] { hd_launch_adi_coef(x.b(), 1, shift, s); });
            launch(x, "adi_solve_uz", [&] { hd_launch_adi_solve(x.b(), 1, shift, s); });
        } else if (x.solver == 1) {
            launch(x, "adi_line_uz", [&] { hd_launch_adi_line(x.b(), 1, shift, s); });
        } else if (x.solver == 3) {
            launch(x, "adi_run_uz", [&] { hd_launch_adi_run(x.b(), 1, shift, x.run_blocks_u, s); });
        } else {
            launch(x, "adi_spec_uz", [&] { hd_launch_adi_spec(x.b(), 1, shift, x.run_blocks_u, own, s); });
        }
        if (x.solver == 0) {
            /* smoothing of t_u, the cells of solveV and the rows' margins by
             * row tiles in one launch */
            launch(x, "smooth_u_explicit_v", [&] { hd_launch_smooth_u_explicit_v(x.b(), x.bw, ntiles, own, s); });
            x.swap_u();
        } else {
            launch(x, "smooth_u", [&] { hd_launch_normalize_post(x.b(), 1, 0, 1, x.bw, ntiles, s); });
            x.swap_u();
            if (x.solver == 3) {
                launch(x, "explicit_margins_v", [&] { hd_launch_explicit_margins(x.b(), 0, 1, s); });
            } else {
                launch(x, "explicit_v", [&] { hd_launch_explicit(x.b(), 0, s); });
                launch(x, "margins_v", [&] { hd_launch_margins(x.b(), 0, s); });
            }
        }
        if (x.solver == 0) {
            /* smoothing of t_v + update_h_moi and the row wetting in one launch */
            launch(x, "smooth_v_wetting_rows", [&] { hd_launch_smooth_wet(x.b(), 0, 1, ntiles, x.wet_mode, s); });
            x.swap_v();
        } else {
            launch(x, "smooth_v_hmoi", [&] { hd_launch_normalize_post(x.b(), 0, 1, 1, x.bw, ntiles, s); });
            x.swap_v();
            launch(x, "wetting_rows", [&] { hd_launch_wetting(x.b(), 1, x.wet_mode, s); });
        }
    } else {
        int shift = (channel && o.kenhhepng == 1) ? 1 : 0;
        slo = 2 + shift;
        if (x.solver == 2) {
            launch(x, "adi_coef_vz", [&] { hd_launch_adi_coef(x.b(), 0, shift, s); });
            launch(x, "adi_solve_vz", [&] { hd_launch_adi_solve(x.b(), 0, shift, s); });
        } else if (x.solver == 1) {
            launch(x, "adi_line_vz", [&] { hd_launch_adi_line(x.b(), 0, shift, s); });
        } else if (x.solver == 3) {
            launch(x, "adi_run_vz", [&] { hd_launch_adi_run(x.b(), 0, shift, x.run_blocks_v, s); });
        } else {
            launch(x, "adi_spec_vz", [&] { hd_launch_adi_spec(x.b(), 0, shift, x.run_blocks_v, own, s); });
        }
        if (x.solver == 0) {
            /* smoothing of t_v, the cells of solveU and the columns' margins
             * by row tiles in one launch */
            launch(x, "smooth_v_explicit_u", [&] { hd_launch_smooth_v_explicit_u(x.b(), x.bw, ntiles, own, s); });
            x.swap_v();
        } else {
            launch(x, "smooth_v", [&] { hd_launch_normalize_post(x.b(), 0, 0, 0, x.bw, ntiles, s); });
            x.swap_v();
            if (x.solver == 3) {
                launch(x, "explicit_margins_u", [&] { hd_launch_explicit_margins(x.b(), 1, 1, s); });
            } else {
                launch(x, "explicit_u", [&] { hd_launch_explicit(x.b(), 1, s); });
                launch(x, "margins_u", [&] { hd_launch_margins(x.b(), 1, s); });
            }
        }
        if (x.solver == 0) {
            /* smoothing of t_u + update_h_moi and the column wetting in one launch */
            launch(x, "smooth_u_wetting_cols", [&] { hd_launch_smooth_wet(x.b(), 1, 0, ntiles, x.wet_mode, s); });
            x.swap_u();
        } else {
            launch(x, "smooth_u_hmoi", [&] { hd_launch_normalize_post(x.b(), 1, 1, 0, x.bw, ntiles, s); });
            x.swap_u();
            launch(x, "wetting_cols", [&] { hd_launch_wetting(x.b(), 0, x.wet_mode, s); });
        }
    }
    if (x.solver == 0) {
        launch(x, "uvz_segments", [&] { hd_launch_uvz_segments(x.b(), o.kenhhepd, o.kenhhepng, ntiles, count_wet, s); });
    } else {
        launch(x, "uvz_dry_htd", [&] { hd_launch_uvz_dry(x.b(), o.kenhhepd, o.kenhhepng, x.bw, ntiles, s); });
        launch(x, "segments_fs_edges", [&] { hd_launch_segments(x.b(), count_wet, s); });
    }
    if (sediment) {
        launch(x, "vth_hesok", [&] { hd_launch_vth_hesok(x.b(), sed_ntiles, s); });
        if (x.solver_v1) {
            launch(x, which == 0 ? "fs_coef_j" : "fs_coef_i",
                   [&] { hd_launch_fs_coef(x.b(), which, o.bed_change_start, o.ketdinh, s); });
            launch(x, which == 0 ? "fs_solve_j" : "fs_solve_i", [&] { hd_launch_fs_solve(x.b(), which, s); });
        } else if (x.solver == 1) {
            launch(x, which == 0 ? "fs_line_j" : "fs_line_i",
                   [&] { hd_launch_fs_line(x.b(), which, o.bed_change_start, o.ketdinh, s); });
        } else {
            launch(x, which == 0 ? (x.solver == 0 ? "fs_spec_j" : "fs_run_j") : (x.solver == 0 ? "fs_spec_i" : "fs_run_i"), [&] {
                hd_launch_fs_run(x.b(), which, o.bed_change_start, o.ketdinh,
                                 which == 0 ? x.run_blocks_u : x.run_blocks_v, x.solver == 0, s);
            });
        }
        launch(x, "update_fs", [&] { hd_launch_update_fs(x.b(), sed_ntiles, s); });
        if (which == 1) { slo = 3; eoff = 1; }
    }
    if (slo_out) *slo_out = slo;
    if (eoff_out) *eoff_out = eoff;
}

static void enqueue_bed(Exec& x, const hd_step_opts& o, int slo, int eoff, int bed_ntiles)
{
    launch(x, "calc_qb", [&] { hd_launch_calc_qb(x.b(), o.qb_mode, bed_ntiles, x.stream); });
    launch(x, "bedload", [&] { hd_launch_bedload(x.b(), o.ketdinh, slo, eoff, bed_ntiles, x.stream); });
}

static void enqueue_step(Exec& x, const hd_step_opts& o, const GraphKey& k)
{
    int slo = 2, eoff = 0;
    enqueue_half(x, 0, o, k.sed0, 0, k.ntiles, k.sed_ntiles, nullptr, nullptr);
    enqueue_half(x, 1, o, k.sed1, k.count, k.ntiles, k.sed_ntiles, &slo, &eoff);
    if (k.bed) enqueue_bed(x, o, slo, eoff, k.bed_ntiles);
}

static GraphKey make_key(hd_ctx* c, double t, const hd_step_opts& o, int count)
{
    GraphKey k;
    double t1 = t + c->P.dT * 0.5;
    double t2 = t1 + c->P.dT * 0.5;
    k.sed0 = t1 >= o.sediment_start;
    k.sed1 = t2 >= o.sediment_start;
    k.bed = bed_trigger(t2, o);
    k.count = count;
    k.kd = o.kenhhepd; k.kn = o.kenhhepng;
    k.analytic = analytic_mode(o);
    k.ketdinh = o.ketdinh; k.qb_mode = o.qb_mode;
    k.s_start = o.bed_change_start;
    k.ntiles = c->settled ? c->ntiles : 0;
    /* the first step that runs the sediment sweep (the bed-load flux) does
     * it over the whole grid; later ones over the active tiles */
    k.sed_ntiles = (c->settled && c->sed_settled) ? c->ntiles : 0;
    k.bed_ntiles = (c->settled && c->bed_settled) ? c->ntiles : 0;
    if (k.sed0 || k.sed1) c->sed_settled = true;
    if (k.bed) c->bed_settled = true;
    return k;
}

static void set_time(const Exec& x, double t)
{
    hipLaunchKernelGGL(k_set_t, dim3(1, 1, x.n), dim3(64), 0, x.stream, x.dv, t);
}

int hd_step(hd_ctx* c, double t, const hd_step_opts* o, double* t_out)
{
    if (!c || !o) return HD_EINVAL;
    if (!c->initialised) return fail(c, HD_ESTATE, "hd_step before hd_init");
    HIPCHK(c, hipSetDevice(c->device));
    double t1 = t + c->P.dT * 0.5, t2 = t1 + c->P.dT * 0.5;
    int rc;
    if ((rc = check_series(c, t1, *o)) || (rc = check_series(c, t2, *o))) return rc;
    Exec x = exec_of(c);
    set_time(x, t);
    GraphKey k = make_key(c, t, *o, 0);
    enqueue_step(x, *o, k);
    HIPCHK(c, hipGetLastError());
    c->settled = true;
    if (t_out) *t_out = t2;
    return HD_OK;
}

int hd_halfstep(hd_ctx* c, int which, double t, const hd_step_opts* o)
{
    if (!c || !o || (which != 0 && which != 1)) return HD_EINVAL;
    if (!c->initialised) return fail(c, HD_ESTATE, "hd_halfstep before hd_init");
    HIPCHK(c, hipSetDevice(c->device));
    int rc = check_series(c, t, *o);
    if (rc) return rc;
    Exec x = exec_of(c);
    set_time(x, t - c->P.dT * 0.5);
    enqueue_half(x, which, *o, t >= o->sediment_start, 0, 0, 0, nullptr, nullptr);
    HIPCHK(c, hipGetLastError());
    return HD_OK;
}

int hd_bed_change(hd_ctx* c, double t, const hd_step_opts* o)
{
    if (!c || !o) return HD_EINVAL;
    int slo = 2, eoff = 0;
    if ((o->kenhhepd || o->kenhhepng) && o->kenhhepng == 1) slo = 3;
    if (t >= o->sediment_start) { slo = 3; eoff = 1; }
    Exec x = exec_of(c);
    enqueue_bed(x, *o, slo, eoff, 0);
    HIPCHK(c, hipGetLastError());
    return HD_OK;
}

/* Enqueue nsteps full steps from t: the first time a step kind (trigger
 * pattern) occurs it is captured into a HIP graph, later steps replay it.
 * A full step swaps each smoothing pair twice, so the replay's pointers
 * (the state-0 device copies) hold.  Returns 0 or the HIP error. */
typedef void (*EnqueueFn)(Exec&, const hd_step_opts&, const GraphKey&);

static hipError_t run_steps(Exec& x, std::map<GraphKey, hipGraphExec_t>& graphs, double& t, int nsteps,
                            const hd_step_opts& o, double dT, int count, const std::function<GraphKey(double)>& key,
                            std::string& why, EnqueueFn enq = enqueue_step)
{
    for (int n = 0; n < nsteps; ++n) {
        GraphKey k = key(t);
        double t2 = t + dT * 0.5 + dT * 0.5;
        if (x.prof && x.prof->on) {
            enq(x, o, k);
            t = t2;
            continue;
        }
        auto it = graphs.find(k);
        if (it == graphs.end()) {
            int st0 = x.state;
            hipGraph_t g;
            hipError_t e = hipStreamBeginCapture(x.stream, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) { why = "begin capture"; return e; }
            enq(x, o, k);
            e = hipStreamEndCapture(x.stream, &g);
            if (e != hipSuccess) { why = "graph capture"; return e; }
            hipGraphExec_t ge;
            e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) { why = "graph instantiate"; return e; }
            if (x.state != st0) {
                hipGraphExecDestroy(ge);
                why = "step changed the buffer assignment";
                return hipErrorUnknown;
            }
            it = graphs.emplace(k, ge).first;
        }
        hipError_t e = hipGraphLaunch(it->second, x.stream);
        if (e != hipSuccess) { why = "graph launch"; return e; }
        t = t2;
    }
    (void)count;
    return hipSuccess;
}

/* The sediment half of one step on a frozen flow: per half-step the
 * boundary values (device time advanced) and, when the sediment module is on,
 * VTH / hesoK, the FS sweep of that direction and the FS update; then the
 * bed change when triggered. */
static void enqueue_sed_step(Exec& x, const hd_step_opts& o, const GraphKey& k)
{
    int an = analytic_mode(o) ? 1 : 0;
    int slo = 2, eoff = 0;
    for (int which = 0; which < 2; ++which) {
        launch(x, "boundary", [&] { hd_launch_boundary_t(x.b(), an, x.stream); });
        if (which == 0 ? k.sed0 : k.sed1) {
            launch(x, "vth_hesok", [&] { hd_launch_vth_hesok(x.b(), k.sed_ntiles, x.stream); });
            launch(x, which == 0 ? (x.solver == 0 ? "fs_spec_j" : "fs_run_j") : (x.solver == 0 ? "fs_spec_i" : "fs_run_i"), [&] {
                hd_launch_fs_run(x.b(), which, o.bed_change_start, o.ketdinh,
                                 which == 0 ? x.run_blocks_u : x.run_blocks_v, x.solver == 0, x.stream);
            });
            launch(x, "update_fs", [&] { hd_launch_update_fs(x.b(), k.sed_ntiles, x.stream); });
            if (which == 1) { slo = 3; eoff = 1; }
        }
    }
    if (k.bed) enqueue_bed(x, o, slo, eoff, k.bed_ntiles);
}

static int run_impl(hd_ctx* c, double t, int nsteps, const hd_step_opts* o, double* t_out, int count)
{
    if (!c || !o || nsteps < 0) return HD_EINVAL;
    if (!c->initialised) return fail(c, HD_ESTATE, "hd_run before hd_init");
    HIPCHK(c, hipSetDevice(c->device));
    for (int n = 0; n < nsteps; ++n) {  /* series long enough for every half-step */
        double tn = t + n * c->P.dT;
        int rc;
        if ((rc = check_series(c, tn + c->P.dT * 0.5, *o)) || (rc = check_series(c, tn + c->P.dT, *o))) return rc;
    }
    Exec x = exec_of(c);
    set_time(x, t);
    std::string why;
    auto key = [&](double tt) { return make_key(c, tt, *o, count); };
    int first = (!c->settled && nsteps > 0) ? 1 : 0;  /* one full-grid step, then the active tiles */
    hipError_t e = run_steps(x, c->graphs, t, first, *o, c->P.dT, count, key, why);
    if (e == hipSuccess && first) c->settled = true;
    if (e == hipSuccess) e = run_steps(x, c->graphs, t, nsteps - first, *o, c->P.dT, count, key, why);
    if (e != hipSuccess) return fail(c, HD_EHIP, why + ": " + hipGetErrorString(e));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (t_out) *t_out = t;
    return check_flags(c);
}

int hd_run(hd_ctx* c, double t, int nsteps, const hd_step_opts* o, double* t_out)
{
    return run_impl(c, t, nsteps, o, t_out, 0);
}

/* hd_run that also accumulates the wet-cell count of every step's end state
 * (the start state of the next step) into the counter read by
 * hd_wet_cell_steps. */
int hd_run_counted(hd_ctx* c, double t, int nsteps, const hd_step_opts* o, double* t_out)
{
    return run_impl(c, t, nsteps, o, t_out, 1);
}

int hd_sediment_run(hd_ctx* c, double t, int nsteps, const hd_step_opts* o, double* t_out)
{
    if (!c || !o || nsteps < 0) return HD_EINVAL;
    if (!c->initialised) return fail(c, HD_ESTATE, "hd_sediment_run before hd_init");
    if (c->state() != 0) return fail(c, HD_ESTATE, "hd_sediment_run between half-steps");
    HIPCHK(c, hipSetDevice(c->device));
    for (int n = 0; n < nsteps; ++n) {
        double tn = t + n * c->P.dT;
        int rc;
        if ((rc = check_series(c, tn + c->P.dT * 0.5, *o)) || (rc = check_series(c, tn + c->P.dT, *o))) return rc;
    }
    Exec x = exec_of(c);
    set_time(x, t);
    std::string why;
    hipError_t e = run_steps(x, c->sed_graphs, t, nsteps, *o, c->P.dT, 0,
                             [&](double tt) { return make_key(c, tt, *o, 0); }, why, enqueue_sed_step);
    if (e != hipSuccess) return fail(c, HD_EHIP, why + ": " + hipGetErrorString(e));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (t_out) *t_out = t;
    return check_flags(c);
}

int hd_bed_snapshot(hd_ctx* c)
{
    if (!c) return HD_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    size_t F = (size_t)(c->N + 3) * (c->M + 3);
    if (!c->bed_snap) {
        int rc = dalloc(c, (void**)&c->bed_snap, F * 8, false);
        if (rc) return rc;
    }
    HIPCHK(c, hipMemcpyAsync(c->bed_snap, c->d.dH, F * 8, hipMemcpyDeviceToDevice, c->stream));
    return HD_OK;
}

int hd_bed_extrapolate(hd_ctx* c, double factor)
{
    if (!c) return HD_EINVAL;
    if (!c->bed_snap) return fail(c, HD_ESTATE, "hd_bed_extrapolate before hd_bed_snapshot");
    HIPCHK(c, hipSetDevice(c->device));
    hd_launch_bed_extrapolate(exec_of(c).b(), c->bed_snap, factor, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return HD_OK;
}

int hd_z_snapshot(hd_ctx* c)
{
    if (!c) return HD_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    size_t F = (size_t)(c->N + 3) * (c->M + 3);
    if (!c->z_snap) {
        int rc = dalloc(c, (void**)&c->z_snap, (F + 5) * 8, false);
        if (rc) return rc;
    }
    HIPCHK(c, hipMemcpyAsync(c->z_snap, c->d.z, F * 8, hipMemcpyDeviceToDevice, c->stream));
    return HD_OK;
}

int hd_z_check(hd_ctx* c, double bound, double* max_abs_dz, double* frac_over, double* wet_volume,
               double* wet_volume_snap)
{
    if (!c) return HD_EINVAL;
    if (!c->z_snap) return fail(c, HD_ESTATE, "hd_z_check before hd_z_snapshot");
    if (!(bound >= 0)) return fail(c, HD_EINVAL, "hd_z_check: bound must be >= 0");
    HIPCHK(c, hipSetDevice(c->device));
    size_t F = (size_t)(c->N + 3) * (c->M + 3);
    double* res = c->z_snap + F;
    double zero[5] = {0, 0, 0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(res, zero, sizeof(zero), hipMemcpyHostToDevice, c->stream));
    hd_launch_z_change(exec_of(c).b(), c->z_snap, bound, res, c->stream);
    HIPCHK(c, hipGetLastError());
    double out[5];
    HIPCHK(c, hipMemcpyAsync(out, res, sizeof(out), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    /* cell volumes are per unit area: the sums of depths */
    if (max_abs_dz) *max_abs_dz = out[0];
    if (frac_over) *frac_over = out[4] > 0 ? out[3] / out[4] : 0.0;
    if (wet_volume) *wet_volume = out[1];
    if (wet_volume_snap) *wet_volume_snap = out[2];
    return HD_OK;
}

int hd_z_change(hd_ctx* c, double* max_abs_dz, double* wet_volume, double* wet_volume_snap)
{
    return hd_z_check(c, HUGE_VAL, max_abs_dz, nullptr, wet_volume, wet_volume_snap);
}

int hd_flow_refresh(hd_ctx* c)
{
    if (!c) return HD_EINVAL;
    if (!c->initialised) return fail(c, HD_ESTATE, "hd_flow_refresh before hd_init");
    if (c->state() != 0) return fail(c, HD_ESTATE, "hd_flow_refresh between half-steps");
    HIPCHK(c, hipSetDevice(c->device));
    hd_launch_flow_refresh(exec_of(c).b(), c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->settled = false;
    c->sed_settled = false;
    c->bed_settled = false;
    return HD_OK;
}

int hd_copy_state(hd_ctx* dst, hd_ctx* src)
{
    if (!dst || !src) return HD_EINVAL;
    if (dst->N != src->N || dst->M != src->M || dst->SL != src->SL || dst->series_len != src->series_len ||
        dst->total_time != src->total_time || dst->state_allocs.size() != src->state_allocs.size())
        return fail(dst, HD_ESHAPE, "hd_copy_state: contexts of different cases");
    if (src->state() != 0 || dst->state() != 0) return fail(dst, HD_ESTATE, "hd_copy_state between half-steps");
    HIPCHK(dst, hipSetDevice(dst->device));
    HIPCHK(dst, hipStreamSynchronize(src->stream));
    for (size_t k = 0; k < dst->state_allocs.size(); ++k) {
        if (dst->state_allocs[k].second != src->state_allocs[k].second)
            return fail(dst, HD_ESHAPE, "hd_copy_state: layout");
        HIPCHK(dst, hipMemcpyAsync(dst->state_allocs[k].first, src->state_allocs[k].first, src->state_allocs[k].second,
                                   hipMemcpyDeviceToDevice, dst->stream));
    }
    HIPCHK(dst, hipStreamSynchronize(dst->stream));
    dst->ntiles = src->ntiles;
    dst->settled = src->settled;
    dst->sed_settled = src->sed_settled;
    dst->bed_settled = src->bed_settled;
    dst->initialised = src->initialised;
    return HD_OK;
}

int hd_clone(hd_ctx* src, hd_ctx** out)
{
    if (!src || !out) return HD_EINVAL;
    if (!src->initialised) return fail(src, HD_ESTATE, "hd_clone before hd_init");
    if (src->state() != 0) return fail(src, HD_ESTATE, "hd_clone between half-steps");
    hd_params p = src->P;
    int rc = hd_create(src->device, &p, src->N, src->M, src->series_len, src->total_time, out);
    if (rc) return fail(src, rc, std::string("hd_clone: ") + g_err);
    hd_ctx* c = *out;
    HIPCHK(src, hipStreamSynchronize(src->stream));
    if (c->state_allocs.size() != src->state_allocs.size()) { hd_destroy(c); return fail(src, HD_EHIP, "layout"); }
    for (size_t k = 0; k < c->state_allocs.size(); ++k) {
        if (c->state_allocs[k].second != src->state_allocs[k].second) { hd_destroy(c); return fail(src, HD_EHIP, "layout"); }
        HIPCHK(src, hipMemcpyAsync(c->state_allocs[k].first, src->state_allocs[k].first, src->state_allocs[k].second,
                                   hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHK(src, hipStreamSynchronize(c->stream));
    set_knobs(c, src->rb_mult, src->bw);
    c->ntiles = src->ntiles;
    c->settled = src->settled;
    c->sed_settled = src->sed_settled;
    c->bed_settled = src->bed_settled;
    c->initialised = true;
    return HD_OK;
}

/* Time `steps` replayed steps of a context from t (events on its stream). */
static int time_steps(hd_ctx* c, double t, const hd_step_opts* o, int steps, double* ms)
{
    int rc = run_impl(c, t, 2, o, nullptr, 0); /* capture */
    if (rc) return rc;
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    HIPCHK(c, hipEventCreate(&b));
    HIPCHK(c, hipEventRecord(a, c->stream));
    rc = run_impl(c, t + 2 * c->P.dT, steps, o, nullptr, 0);
    HIPCHK(c, hipEventRecord(b, c->stream));
    HIPCHK(c, hipEventSynchronize(b));
    float f = 0;
    HIPCHK(c, hipEventElapsedTime(&f, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    *ms = f / steps;
    return rc;
}

int hd_autotune(hd_ctx* c, double t, const hd_step_opts* o, int steps, int* rb_mult, int* bw, double* ms_per_step)
{
    if (!c || !o || steps < 1) return HD_EINVAL;
    static const int kRb[] = {1, 2, 4};
    static const int kBw[] = {64, 128, 256};
    double best = 1e30;
    int brb = c->rb_mult, bbw = c->bw;
    for (int r : kRb)
        for (int w : kBw) {
            hd_ctx* s = nullptr;
            int rc = hd_clone(c, &s);
            if (rc) return rc;
            set_knobs(s, r, w);
            double ms = 0;
            rc = time_steps(s, t, o, steps, &ms);
            if (rc) {
                std::string m = s->err;
                hd_destroy(s);
                return fail(c, rc, "autotune: " + m);
            }
            hd_destroy(s);
            if (ms < best) { best = ms; brb = r; bbw = w; }
        }
    for (auto& g : c->graphs) hipGraphExecDestroy(g.second);
    c->graphs.clear();
    for (auto& g : c->sed_graphs) hipGraphExecDestroy(g.second);
    c->sed_graphs.clear();
    set_knobs(c, brb, bbw);
    if (rb_mult) *rb_mult = brb;
    if (bw) *bw = bbw;
    if (ms_per_step) *ms_per_step = best;
    return HD_OK;
}

int hd_set_launch(hd_ctx* c, int rb_mult, int bw)
{
    if (!c || rb_mult < 1 || rb_mult > 16 || (bw != 64 && bw != 128 && bw != 256)) return HD_EINVAL;
    for (auto& g : c->graphs) hipGraphExecDestroy(g.second);
    c->graphs.clear();
    for (auto& g : c->sed_graphs) hipGraphExecDestroy(g.second);
    c->sed_graphs.clear();
    set_knobs(c, rb_mult, bw);
    return HD_OK;
}

int hd_reset_wet_counter(hd_ctx* c)
{
    if (!c) return HD_EINVAL;
    unsigned long long z = 0;
    HIPCHK(c, hipMemcpyAsync(c->d.wet_count, &z, 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return HD_OK;
}

int hd_wet_cell_steps(hd_ctx* c, long long* out)
{
    if (!c || !out) return HD_EINVAL;
    HIPCHK(c, hipMemcpyAsync(out, c->d.wet_count, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return HD_OK;
}

int hd_wet_count(hd_ctx* c, long long* out)
{
    if (!c || !out) return HD_EINVAL;
    unsigned long long z = 0;
    /* slot 1: the one-shot count (slot 0 is the run accumulator) */
    HIPCHK(c, hipMemcpyAsync(c->d.wet_count + 1, &z, 8, hipMemcpyHostToDevice, c->stream));
    hd_launch_wet_count(exec_of(c).b(), c->stream);
    HIPCHK(c, hipMemcpyAsync(out, c->d.wet_count + 1, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return HD_OK;
}

int hd_sync(hd_ctx* c)
{
    if (!c) return HD_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return check_flags(c);
}

int hd_struct_ptrs(hd_ctx* c, void** arg, void** arr)
{
    if (!c) return HD_EINVAL;
    const HdDev& d = c->d;
    ArgMirror a;
    std::memset(&a, 0, sizeof(a));
    a.M = c->M; a.N = c->N;
    double hm[4];
    int rc = hd_get_hmax(c, hm);
    if (rc) return rc;
    a.hmax_u = hm[0]; a.hmax_d = hm[1]; a.hmax_l = hm[2]; a.hmax_r = hm[3];
    a.bienQ = d.bienQ; a.daui = d.daui; a.dauj = d.dauj; a.cuoii = d.cuoii; a.cuoij = d.cuoij;
    a.moci = d.moci; a.mocj = d.mocj; a.khouot = d.khouot; a.boundary_type = d.boundary_type;
    a.h = d.h; a.v = d.v; a.u = d.u; a.z = d.z; a.t_u = d.t_u; a.t_v = d.t_v; a.t_z = d.t_z;
    a.Htdu = d.Htdu; a.Htdv = d.Htdv; a.H_moi = d.H_moi; a.htaiz = d.htaiz; a.htaiz_bd = d.htaiz_bd;
    a.ubt = d.ubt; a.ubp = d.ubp; a.vbt = d.vbt; a.vbd = d.vbd;
    a.hsnham = d.hsnham; a.VISCOIDX = d.VISCOIDX; a.Kx1 = d.Kx1; a.Ky1 = d.Ky1; a.Tsyw = d.Tsyw; a.Tsxw = d.Tsxw;
    a.bc_up = d.bc_up; a.bc_down = d.bc_down; a.bc_left = d.bc_left; a.bc_right = d.bc_right; a.hi = d.hi;
    a.FS = d.FS; a.tFS = d.tFS; a.CC_u = d.CC_u; a.CC_d = d.CC_d; a.CC_l = d.CC_l; a.CC_r = d.CC_r;
    a.VTH = d.VTH; a.Kx = d.Kx; a.Ky = d.Ky; a.Fw = d.Fw; a.Qbx = d.Qbx; a.Qby = d.Qby; a.dH = d.dH;
    ArrMirror r;
    std::memset(&r, 0, sizeof(r));
    r.AA = d.AA; r.BB = d.BB; r.CC = d.CC; r.DD = d.DD; r.Ap = d.Ap; r.Bp = d.Bp; r.SN = d.seg_sn;
    HIPCHK(c, hipMemcpyAsync(c->arg_mirror, &a, sizeof(a), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->arr_mirror, &r, sizeof(r), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (arg) *arg = c->arg_mirror;
    if (arr) *arr = c->arr_mirror;
    return HD_OK;
}

static void prof_reset(Prof& p, bool on)
{
    for (auto& r : p.pending) { p.pool.push_back(r.a); p.pool.push_back(r.b); }
    p.pending.clear();
    p.acc.clear();
    p.order.clear();
    p.on = on;
}

static hipError_t prof_collect(Prof& p, int n, const char** names, double* total_ms, int* launches, int* count)
{
    for (auto& r : p.pending) {
        float ms = 0.f;
        hipError_t e = hipEventElapsedTime(&ms, r.a, r.b);
        if (e != hipSuccess) return e;
        auto it = p.acc.find(r.name);
        if (it == p.acc.end()) { p.order.push_back(r.name); it = p.acc.emplace(r.name, std::make_pair(0.0, 0)).first; }
        it->second.first += ms;
        it->second.second += 1;
        p.pool.push_back(r.a);
        p.pool.push_back(r.b);
    }
    p.pending.clear();
    int k = 0;
    for (const std::string& nm : p.order) {
        if (k >= n) break;
        auto it = p.acc.find(nm);
        if (names) names[k] = it->first.c_str();
        if (total_ms) total_ms[k] = it->second.first;
        if (launches) launches[k] = it->second.second;
        ++k;
    }
    *count = k;
    return hipSuccess;
}

int hd_profile(hd_ctx* c, int enable)
{
    if (!c) return HD_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    prof_reset(c->prof, enable != 0);
    return HD_OK;
}

int hd_kernel_times(hd_ctx* c, int n, const char** names, double* total_ms, int* launches)
{
    if (!c) return HD_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int k = 0;
    HIPCHK(c, prof_collect(c->prof, n, names, total_ms, launches, &k));
    return k;
}

/* ================================================================ ensembles */

struct hd_ensemble {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<hd_ctx*> m;
    HdDev* dv = nullptr;   /* [4][n] device copies, state-major */
    int N = 0, M = 0, run_blocks_u = 0, run_blocks_v = 0;
    std::map<GraphKey, hipGraphExec_t> graphs;
    Prof prof;
    std::string err;
};

static int efail(hd_ensemble* e, int code, const std::string& msg)
{
    if (e) e->err = msg;
    g_err = msg;
    return code;
}

int hd_ensemble_create(hd_ctx* const* members, int n, hd_ensemble** out)
{
    if (!members || n < 1 || !out) return efail(nullptr, HD_EINVAL, "hd_ensemble_create: bad arguments");
    hd_ctx* c0 = members[0];
    for (int k = 0; k < n; ++k) {
        hd_ctx* c = members[k];
        if (!c || !c->initialised) return efail(nullptr, HD_ESTATE, "ensemble member not initialised");
        if (c->device != c0->device || c->P.dT != c0->P.dT || c->solver != c0->solver ||
            c->solver_v1 != c0->solver_v1)
            return efail(nullptr, HD_EINVAL, "ensemble members must share device, dT and solver form");
        if (c->state() != 0) return efail(nullptr, HD_ESTATE, "ensemble member between half-steps");
    }
    if (hipSetDevice(c0->device) != hipSuccess) return efail(nullptr, HD_EHIP, "hipSetDevice");
    hd_ensemble* e = new hd_ensemble();
    e->device = c0->device;
    e->m.assign(members, members + n);
    for (hd_ctx* c : e->m) {
        if (hipStreamSynchronize(c->stream) != hipSuccess) { delete e; return efail(nullptr, HD_EHIP, "sync"); }
        e->N = std::max(e->N, c->N);
        e->M = std::max(e->M, c->M);
        e->run_blocks_u = std::max(e->run_blocks_u, c->run_blocks_u);
        e->run_blocks_v = std::max(e->run_blocks_v, c->run_blocks_v);
    }
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return efail(nullptr, HD_EHIP, "hipStreamCreate");
    }
    std::vector<HdDev> st((size_t)4 * n);
    for (int s = 0; s < 4; ++s)
        for (int k = 0; k < n; ++k) {
            HdDev d = e->m[k]->d;
            if (s & 1) std::swap(d.t_u, d.tmp_u);
            if (s & 2) std::swap(d.t_v, d.tmp_v);
            st[(size_t)s * n + k] = d;
        }
    if (hipMalloc((void**)&e->dv, st.size() * sizeof(HdDev)) != hipSuccess ||
        hipMemcpy(e->dv, st.data(), st.size() * sizeof(HdDev), hipMemcpyHostToDevice) != hipSuccess) {
        hd_ensemble_destroy(e);
        return efail(nullptr, HD_ENOMEM, "ensemble device tables");
    }
    *out = e;
    return HD_OK;
}

int hd_ensemble_destroy(hd_ensemble* e)
{
    if (!e) return HD_OK;
    hipSetDevice(e->device);
    if (e->stream) hipStreamSynchronize(e->stream);
    for (auto& g : e->graphs) hipGraphExecDestroy(g.second);
    for (auto& r : e->prof.pending) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
    for (auto ev : e->prof.pool) hipEventDestroy(ev);
    if (e->dv) hipFree(e->dv);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
    return HD_OK;
}

const char* hd_ensemble_last_error(const hd_ensemble* e) { return e ? e->err.c_str() : g_err.c_str(); }

int hd_ensemble_run(hd_ensemble* e, double t, int nsteps, const hd_step_opts* o, int counted, double* t_out)
{
    if (!e || !o || nsteps < 0) return efail(e, HD_EINVAL, "hd_ensemble_run: bad arguments");
    if (hipSetDevice(e->device) != hipSuccess) return efail(e, HD_EHIP, "hipSetDevice");
    hd_ctx* c0 = e->m[0];
    double dT = c0->P.dT;
    for (size_t mi = 0; mi < e->m.size(); ++mi) {
        hd_ctx* c = e->m[mi];
        for (int k = 0; k < nsteps; ++k) {
            double tn = t + k * dT;
            int rc;
            if ((rc = check_series(c, tn + dT * 0.5, *o)) || (rc = check_series(c, tn + dT, *o)))
                return efail(e, rc, "member " + std::to_string(mi) + ": " + c->err);
        }
        /* work queued on a member's own stream comes first */
        if (hipStreamSynchronize(c->stream) != hipSuccess) return efail(e, HD_EHIP, "member sync");
    }
    int count = counted ? 1 : 0;
    Exec x{e->stream, e->dv, (int)e->m.size(), e->N, e->M, 0, c0->solver, c0->solver_v1, c0->wet_mode, e->run_blocks_u,
           e->run_blocks_v, c0->bw, &e->prof, nullptr};
    set_time(x, t);
    std::string why;
    bool settled = true, sed_settled = true, bed_settled = true;
    int maxtiles = 0;
    for (hd_ctx* c : e->m) {
        settled = settled && c->settled;
        sed_settled = sed_settled && c->sed_settled;
        bed_settled = bed_settled && c->bed_settled;
        maxtiles = std::max(maxtiles, c->ntiles);
    }
    auto key = [&](double tt) {
        GraphKey k = make_key(c0, tt, *o, count);
        k.ntiles = settled ? maxtiles : 0;
        k.sed_ntiles = (settled && sed_settled) ? maxtiles : 0;
        k.bed_ntiles = (settled && bed_settled) ? maxtiles : 0;
        if (k.sed0 || k.sed1) {
            sed_settled = true;
            for (hd_ctx* c : e->m) c->sed_settled = true;
        }
        if (k.bed) {
            bed_settled = true;
            for (hd_ctx* c : e->m) c->bed_settled = true;
        }
        return k;
    };
    int first = (!settled && nsteps > 0) ? 1 : 0;
    hipError_t err = run_steps(x, e->graphs, t, first, *o, dT, count, key, why);
    if (err == hipSuccess && first) {
        settled = true;
        for (hd_ctx* c : e->m) c->settled = true;
    }
    if (err == hipSuccess) err = run_steps(x, e->graphs, t, nsteps - first, *o, dT, count, key, why);
    if (err != hipSuccess) return efail(e, HD_EHIP, why + ": " + hipGetErrorString(err));
    if ((err = hipGetLastError()) != hipSuccess) return efail(e, HD_EHIP, hipGetErrorString(err));
    if ((err = hipStreamSynchronize(e->stream)) != hipSuccess) return efail(e, HD_EHIP, hipGetErrorString(err));
    for (size_t k = 0; k < e->m.size(); ++k) {
        int rc = check_flags(e->m[k]);
        if (rc) return efail(e, rc, "member " + std::to_string(k) + ": " + e->m[k]->err);
    }
    if (t_out) *t_out = t;
    return HD_OK;
}

int hd_ensemble_profile(hd_ensemble* e, int enable)
{
    if (!e) return HD_EINVAL;
    if (hipStreamSynchronize(e->stream) != hipSuccess) return efail(e, HD_EHIP, "sync");
    prof_reset(e->prof, enable != 0);
    return HD_OK;
}

int hd_ensemble_kernel_times(hd_ensemble* e, int n, const char** names, double* total_ms, int* launches)
{
    if (!e) return HD_EINVAL;
    if (hipStreamSynchronize(e->stream) != hipSuccess) return efail(e, HD_EHIP, "sync");
    int k = 0;
    hipError_t err = prof_collect(e->prof, n, names, total_ms, launches, &k);
    if (err != hipSuccess) return efail(e, HD_EHIP, hipGetErrorString(err));
    return k;
}

static int thomas_batch_impl(int device, int nsys, const int* off, const int* sn, const double* A, const double* B,
                             const double* C, const double* D, double* x, int nrows, bool spec)
{
    if (nsys < 0 || nrows < 0 || !off || !sn || !A || !B || !C || !D || !x) return fail(nullptr, HD_EINVAL, "bad args");
    for (int s = 0; s < nsys; ++s)
        if (off[s] < 0 || (sn[s] >= 0 && off[s] + sn[s] >= nrows)) return fail(nullptr, HD_EINVAL, "system out of range");
    if (hipSetDevice(device) != hipSuccess) return fail(nullptr, HD_EHIP, "hipSetDevice");
    std::vector<void*> bufs;
    auto mk = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return nullptr;
        bufs.push_back(p);
        return p;
    };
    size_t rb = (size_t)nrows * 8;
    int *doff = (int*)mk(nsys * 4), *dsn = (int*)mk(nsys * 4);
    double *dA = (double*)mk(rb), *dB = (double*)mk(rb), *dC = (double*)mk(rb), *dD = (double*)mk(rb);
    double *dx = (double*)mk(rb), *dAp = (double*)mk(rb), *dBp = (double*)mk(rb);
    int rc = HD_OK;
    if (!doff || !dsn || !dA || !dB || !dC || !dD || !dx || !dAp || !dBp) rc = fail(nullptr, HD_ENOMEM, "hipMalloc");
    if (!rc) {
        hipMemcpy(doff, off, nsys * 4, hipMemcpyHostToDevice);
        hipMemcpy(dsn, sn, nsys * 4, hipMemcpyHostToDevice);
        hipMemcpy(dA, A, rb, hipMemcpyHostToDevice);
        hipMemcpy(dB, B, rb, hipMemcpyHostToDevice);
        hipMemcpy(dC, C, rb, hipMemcpyHostToDevice);
        hipMemcpy(dD, D, rb, hipMemcpyHostToDevice);
        hipMemcpy(dx, x, rb, hipMemcpyHostToDevice);
        if (spec) hd_thomas_spec_device(nsys, doff, dsn, dA, dB, dC, dD, dx, dAp, dBp, nrows, 0);
        else hd_thomas_batch_device(nsys, doff, dsn, dA, dB, dC, dD, dx, dAp, dBp, nrows, 0);
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = fail(nullptr, HD_EHIP, hipGetErrorString(e));
        else hipMemcpy(x, dx, rb, hipMemcpyDeviceToHost);
    }
    for (void* p : bufs) hipFree(p);
    return rc;
}

int hd_thomas_batch(int device, int nsys, const int* off, const int* sn, const double* A, const double* B,
                    const double* C, const double* D, double* x, int nrows)
{
    return thomas_batch_impl(device, nsys, off, sn, A, B, C, D, x, nrows, false);
}

int hd_thomas_batch_spec(int device, int nsys, const int* off, const int* sn, const double* A, const double* B,
                         const double* C, const double* D, double* x, int nrows)
{
    return thomas_batch_impl(device, nsys, off, sn, A, B, C, D, x, nrows, true);
}


/* ------------------------------------------------------------ per-kernel entry points */

/* The library's kernel(s) for one reference kernel, on the current state of a
 * context between whole steps, waited for before returning.  For unit parity
 * against the checker; the step entry points are the product path. */
static int unit_begin(hd_ctx* c, Exec& x)
{
    if (!c) return HD_EINVAL;
    if (c->state() != 0) return fail(c, HD_ESTATE, "per-kernel entry point inside a half-step");
    HIPCHK(c, hipSetDevice(c->device));
    x = exec_of(c);
    return HD_OK;
}

static int unit_end(hd_ctx* c, bool mask_changed)
{
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (mask_changed) {
        /* the next step re-derives its active tiles over the whole grid */
        c->settled = false;
        c->sed_settled = false;
        c->bed_settled = false;
    }
    return check_flags(c);
}

/* rows of the line-major buffers (line stride 2*max(N,M)+4) into a caller
 * array with the reference's line stride */
static int rows_out(hd_ctx* c, double* dev, double* host, int nlines, int stride)
{
    if (!host) return HD_OK;
    int L = c->N > c->M ? c->N : c->M;
    size_t tw = 2 * (size_t)L + 4;
    int w = (size_t)stride < tw ? stride : (int)tw;
    HIPCHK(c, hipMemcpy2D(host, (size_t)stride * 8, dev, tw * 8, (size_t)w * 8, nlines, hipMemcpyDeviceToHost));
    return HD_OK;
}

int hd_k_Onetime_init(hd_ctx* c)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    hd_launch_onetime_init(x.b(), x.stream);
    return unit_end(c, true);
}

int hd_k_Find_Calculation_limits(hd_ctx* c)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    hd_launch_find_limits(x.b(), x.stream);
    if ((rc = unit_end(c, true))) return rc;
    return compute_tiles(c);
}

int hd_k_Htuongdoi(hd_ctx* c)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    hd_launch_htuongdoi(x.b(), x.stream);
    return unit_end(c, false);
}

int hd_k_preprocess_data(hd_ctx* c)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    hd_launch_preprocess(x.b(), x.stream);
    return unit_end(c, false);
}

int hd_k_Update_Boundary_Value(hd_ctx* c, double t, int analytic)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    hd_step_opts o{};
    o.boundary_mode = analytic ? HD_BND_ANALYTIC : HD_BND_SERIES;
    if ((rc = check_series(c, t, o))) return rc;
    set_time(x, t - c->P.dT * 0.5);          /* the boundary kernel advances it by dT/2 */
    hd_launch_boundary_t(x.b(), analytic ? 1 : 0, x.stream);
    return unit_end(c, false);
}

static int k_sweep(hd_ctx* c, int isU, int shift)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    hd_launch_adi_spec(x.b(), isU, shift, isU ? x.run_blocks_u : x.run_blocks_v, 0, x.stream);
    return unit_end(c, false);
}

int hd_k_UZSolver(hd_ctx* c, int shift) { return k_sweep(c, 1, shift); }
int hd_k_VZSolver(hd_ctx* c, int shift) { return k_sweep(c, 0, shift); }

static int k_matrix(hd_ctx* c, int isU, int shift, double* AA, double* BB, double* CC, double* DD, int* SN)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    hd_launch_adi_coef(x.b(), isU, shift, x.stream);
    if ((rc = unit_end(c, false))) return rc;
    int nl = (isU ? c->M : c->N) + 3, stride = 2 * (isU ? c->N : c->M) + 1;
    double* dev[] = {c->d.AA, c->d.BB, c->d.CC, c->d.DD};
    double* host[] = {AA, BB, CC, DD};
    for (int k = 0; k < 4; ++k)
        if ((rc = rows_out(c, dev[k], host[k], nl, stride))) return rc;
    int L = c->N > c->M ? c->N : c->M;
    if (SN) HIPCHK(c, hipMemcpy(SN, c->d.seg_sn, (size_t)(L + 3) * c->SL * 4, hipMemcpyDeviceToHost));
    return HD_OK;
}

int hd_k_UZSolver_calculate_matrix_coeff(hd_ctx* c, int shift, double* AA, double* BB, double* CC, double* DD,
                                         int* SN)
{
    return k_matrix(c, 1, shift, AA, BB, CC, DD, SN);
}

int hd_k_VZSolver_calculate_matrix_coeff(hd_ctx* c, int shift, double* AA, double* BB, double* CC, double* DD,
                                         int* SN)
{
    return k_matrix(c, 0, shift, AA, BB, CC, DD, SN);
}

int hd_k_Normalize(hd_ctx* c, int isU)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    hd_launch_normalize_post(x.b(), isU ? 1 : 0, 0, isU ? 0 : 1, x.bw, 0, x.stream);
    size_t F = (size_t)(c->N + 3) * (c->M + 3) * 8;
    /* update_buffer: the smoothed partner back into t_u / t_v */
    HIPCHK(c, hipMemcpyAsync(isU ? c->d.t_u : c->d.t_v, isU ? c->d.tmp_u : c->d.tmp_v, F, hipMemcpyDeviceToDevice,
                             x.stream));
    return unit_end(c, false);
}

static int k_explicit(hd_ctx* c, int isU)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    hd_launch_explicit(x.b(), isU, x.stream);
    hd_launch_margins(x.b(), isU, x.stream);
    return unit_end(c, false);
}

int hd_k_solveV(hd_ctx* c) { return k_explicit(c, 0); }
int hd_k_solveU(hd_ctx* c) { return k_explicit(c, 1); }

int hd_k_Reset_states(hd_ctx* c, int horizontal, int kenhhepd, int kenhhepng)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    /* update_h_moi, the drying test and the wetting candidates (its smoothing
     * output in the partner buffer is discarded), the serial wetting pass,
     * then the drying writes with update_uvz and Htuongdoi */
    hd_launch_normalize_post(x.b(), horizontal ? 0 : 1, 1, horizontal ? 1 : 0, x.bw, 0, x.stream);
    hd_launch_wetting(x.b(), horizontal ? 1 : 0, x.wet_mode, x.stream);
    hd_launch_uvz_dry(x.b(), kenhhepd, kenhhepng, x.bw, 0, x.stream);
    return unit_end(c, true);
}

int hd_k_Find_VTH_hesoK(hd_ctx* c)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    hd_launch_vth_hesok(x.b(), 0, x.stream);
    return unit_end(c, false);
}

static int k_fs(hd_ctx* c, int which, double t, double s_start, int ketdinh)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    set_time(x, t);
    hd_launch_fs_run(x.b(), which, s_start, ketdinh, which == 0 ? x.run_blocks_u : x.run_blocks_v, 1, x.stream);
    hd_launch_update_fs(x.b(), 0, x.stream);
    return unit_end(c, false);
}

int hd_k_FSj_solve(hd_ctx* c, double t, double s_start, int ketdinh) { return k_fs(c, 0, t, s_start, ketdinh); }
int hd_k_FSi_solve(hd_ctx* c, double t, double s_start, int ketdinh) { return k_fs(c, 1, t, s_start, ketdinh); }

static int k_fs_rows(hd_ctx* c, int which, double t, double s_start, int ketdinh, double* AA, double* BB,
                     double* CC, double* DD)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    set_time(x, t);
    hd_launch_fs_coef(x.b(), which, s_start, ketdinh, x.stream);
    if ((rc = unit_end(c, false))) return rc;
    int nl = (which == 0 ? c->M : c->N) + 3, stride = (which == 0 ? c->N : c->M) + 3;
    double* dev[] = {c->d.AA, c->d.BB, c->d.CC, c->d.DD};
    double* host[] = {AA, BB, CC, DD};
    for (int k = 0; k < 4; ++k)
        if ((rc = rows_out(c, dev[k], host[k], nl, stride))) return rc;
    return HD_OK;
}

int hd_k_Scan_FSj(hd_ctx* c, double t, double s_start, int ketdinh, double* AA, double* BB, double* CC, double* DD)
{
    return k_fs_rows(c, 0, t, s_start, ketdinh, AA, BB, CC, DD);
}

int hd_k_Scan_FSi(hd_ctx* c, double t, double s_start, int ketdinh, double* AA, double* BB, double* CC, double* DD)
{
    return k_fs_rows(c, 1, t, s_start, ketdinh, AA, BB, CC, DD);
}

int hd_k_Calculate_Qb(hd_ctx* c, int qb_mode)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    hd_launch_calc_qb(x.b(), qb_mode, 0, x.stream);
    return unit_end(c, false);
}

int hd_k_BedLoad(hd_ctx* c, int ketdinh, int sidx, int eidx)
{
    Exec x; int rc = unit_begin(c, x); if (rc) return rc;
    if (sidx < 2 || eidx > c->N || sidx > eidx) return fail(c, HD_EINVAL, "hd_k_BedLoad: rows out of range");
    hd_launch_bedload(x.b(), ketdinh, sidx, c->N - eidx, 0, x.stream);
    return unit_end(c, false);
}

/* ------------------------------------------------ batched Thomas, no allocation */

/* Elimination-factor scratch of hd_thomas_batch_noalloc: Ap in the first
 * HD_THOMAS_NOALLOC_MAX_ROWS doubles, Bp in the second (16 MiB per device,
 * held until the process ends).  Host code never uses the symbol's name as
 * an address: each device's copy is reached through the address that
 * hipGetSymbolAddress returned on that device (g_scratch_ptr). */
static __device__ __align__(256) double scratch[2 * (size_t)HD_THOMAS_NOALLOC_MAX_ROWS];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[64];

int hd_thomas_batch_noalloc(int device, int nsys, const int* off, const int* sn, const double* A, const double* B,
                            const double* C, const double* D, double* x, int nrows, void* stream)
{
    if (nsys < 0 || nrows < 0 || (nsys > 0 && (!off || !sn || !A || !B || !C || !D || !x)))
        return fail(nullptr, HD_EINVAL, "hd_thomas_batch_noalloc: bad arguments");
    if (nrows > HD_THOMAS_NOALLOC_MAX_ROWS)
        return fail(nullptr, HD_EINVAL, "hd_thomas_batch_noalloc: nrows " + std::to_string(nrows) +
                                            " exceeds the scratch limit of " +
                                            std::to_string(HD_THOMAS_NOALLOC_MAX_ROWS) + " rows");
    if (device < 0 || device >= 64) return fail(nullptr, HD_EINVAL, "hd_thomas_batch_noalloc: bad device");
    if (hipSetDevice(device) != hipSuccess) return fail(nullptr, HD_EHIP, "hipSetDevice");
    double* base;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[device]) {
            void* p = nullptr;
            hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return fail(nullptr, HD_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            g_scratch_ptr[device] = (double*)p;
        }
        base = g_scratch_ptr[device];
    }
    hipStream_t s = (hipStream_t)stream;
    hd_thomas_batch_device(nsys, off, sn, A, B, C, D, x, base, base + HD_THOMAS_NOALLOC_MAX_ROWS, nrows, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(nullptr, HD_EHIP, std::string("hd_thomas_batch_noalloc: ") + hipGetErrorString(e));
    return HD_OK;
}

} // extern "C"
