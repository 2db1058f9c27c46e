/*
 * hd_api.hip -- the extern "C" boundary of libhydist.so (include/hydist.h):
 * context and device memory, named field transfer, the step sequence of
 * Engine.hydraulic_Calculation (Engine.py:144-366) and its HIP-graph replay.
 */
#include "hd_internal.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

namespace {

__global__ void k_set_t(double* tdev, double t) { tdev[0] = t; }

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
    bool operator<(const GraphKey& o) const
    {
        return std::tie(sed0, sed1, bed, count, kd, kn, analytic, ketdinh, qb_mode, s_start) <
               std::tie(o.sed0, o.sed1, o.bed, o.count, o.kd, o.kn, o.analytic, o.ketdinh, o.qb_mode, o.s_start);
    }
};

const char* kNames[] = {"h", "hsnham", "VISCOIDX", "u", "v", "z", "t_u", "t_v", "t_z", "Htdu", "Htdv", "H_moi",
                        "htaiz", "htaiz_bd", "Kx1", "Ky1", "Tsxw", "Tsyw", "FS", "tFS", "VTH", "Kx", "Ky", "Fw",
                        "Qbx", "Qby", "dH"};

} // namespace

struct hd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int N = 0, M = 0, SL = 0, series_len = 0, total_time = 0;
    hd_params P;
    HdDev d;
    std::map<std::string, Field> fields;
    std::vector<void*> allocs;
    double* tdev = nullptr;
    int* cand_cnt = nullptr;
    int* cand = nullptr;
    unsigned char* dryflag = nullptr;
    void* arg_mirror = nullptr;
    void* arr_mirror = nullptr;
    bool initialised = false;
    std::string err;
    std::map<GraphKey, hipGraphExec_t> graphs;
    /* hydrodynamic solver form (HYDIST_SOLVER): 0 = one wave per wet run
     * (default), 1 = one wave per line ("line"), 2 = rows kernel + one lane
     * per line ("rows").  All three give identical bits. */
    int solver = 0;
    bool solver_v1 = false;     /* two-kernel sediment form as well (HYDIST_SOLVER=rows) */
    int run_blocks_u = 0, run_blocks_v = 0;
    /* per-kernel HIP-event timing (hd_profile / hd_kernel_times) */
    bool profile = false;
    struct Rec { const char* name; hipEvent_t a, b; };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    std::vector<std::string> order;
    std::map<std::string, std::pair<double, int>> acc;
};

static hipEvent_t take_event(hd_ctx* c)
{
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

/* Run one launcher, bracketed by events on the context's stream when
 * profiling is on. */
template <class F>
static void launch(hd_ctx* c, const char* name, F&& f)
{
    if (!c->profile) { f(); return; }
    hd_ctx::Rec r{name, take_event(c), take_event(c)};
    (void)hipEventRecord(r.a, c->stream);
    f();
    (void)hipEventRecord(r.b, c->stream);
    c->pending.push_back(r);
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

static int dalloc(hd_ctx* c, void** p, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(c, HD_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->allocs.push_back(*p);
    e = hipMemset(*p, 0, bytes);
    if (e != hipSuccess) return fail(c, HD_EHIP, std::string("hipMemset: ") + hipGetErrorString(e));
    return HD_OK;
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
        c->solver = c->solver_v1 ? 2 : (sv && std::string(sv) == "line") ? 1 : 0;
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
    c->run_blocks_u = 2 * (M + 3) < (M + 3) * SL ? 2 * (M + 3) : (M + 3) * SL;
    c->run_blocks_v = 2 * (N + 3) < (N + 3) * SL ? 2 * (N + 3) : (N + 3) * SL;
    if ((rc = dalloc(c, (void**)&d.flags, 16 * 4))) { hd_destroy(c); return rc; }
    /* [0]: per-step accumulator of hd_run_counted, [1]: one-shot count of hd_wet_count */
    if ((rc = dalloc(c, (void**)&d.wet_count, 16))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&c->tdev, 8))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&c->cand_cnt, nlines * 4))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&c->cand, nlines * 64 * 4))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, (void**)&c->dryflag, F))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, &c->arg_mirror, sizeof(ArgMirror)))) { hd_destroy(c); return rc; }
    if ((rc = dalloc(c, &c->arr_mirror, sizeof(ArrMirror)))) { hd_destroy(c); return rc; }
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
    for (auto& r : c->pending) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
    for (auto e : c->pool) hipEventDestroy(e);
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

/* rep.py:126-139: Onetime_init, Find_Calculation_limits_x/y, Htuongdoi,
 * preprocess_data. */
int hd_init(hd_ctx* c)
{
    if (!c) return HD_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    const HdDev& d = c->d;
    hd_launch_onetime_init(d, c->stream);
    hd_launch_find_limits(d, c->stream);
    hd_launch_htuongdoi(d, c->stream);
    hd_launch_preprocess(d, c->stream);
    HIPCHK(c, hipGetLastError());
    int rc = check_flags(c);
    if (rc) return rc;
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
    hd_launch_find_limits(c->d, c->stream);
    hd_launch_htuongdoi(c->d, c->stream);
    hd_launch_preprocess(c->d, c->stream);
    HIPCHK(c, hipGetLastError());
    return check_flags(c);
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

static void swap_u(hd_ctx* c) { std::swap(c->d.t_u, c->d.tmp_u); }
static void swap_v(hd_ctx* c) { std::swap(c->d.t_v, c->d.tmp_v); }

/* One half-step of Engine.py:146-239 (which = 0, UZ) or :260-361 (which = 1,
 * VZ), enqueued on the stream.  The device time is advanced by the boundary
 * kernel. */
static void enqueue_half(hd_ctx* c, int which, const hd_step_opts& o, bool sediment, int count_wet, int* sidx_out,
                         int* eidx_out)
{
    hipStream_t s = c->stream;
    int N = c->N, M = c->M;
    bool channel = o.kenhhepd || o.kenhhepng;
    int an = analytic_mode(o) ? 1 : 0;
    launch(c, "boundary", [&] { hd_launch_boundary_t(c->d, c->tdev, an, s); });
    int sidx, eidx;
    if (which == 0) {
        sidx = 2; eidx = M;
        if (channel && o.kenhhepd == 1) { sidx = 3; eidx = M - 1; }
        if (c->solver == 2) {
            launch(c, "adi_coef_uz", [&] { hd_launch_adi_coef(c->d, 1, sidx, eidx, s); });
            launch(c, "adi_solve_uz", [&] { hd_launch_adi_solve(c->d, 1, sidx, eidx, s); });
        } else if (c->solver == 1) {
            launch(c, "adi_line_uz", [&] { hd_launch_adi_line(c->d, 1, sidx, eidx, s); });
        } else {
            launch(c, "adi_run_uz", [&] { hd_launch_adi_run(c->d, 1, sidx, eidx, c->run_blocks_u, s); });
        }
        launch(c, "smooth_u", [&] { hd_launch_normalize_post(c->d, 1, 0, 1, c->cand_cnt, c->cand, c->dryflag, s); });
        swap_u(c);
        launch(c, "explicit_v", [&] { hd_launch_explicit(c->d, 0, 2, N, s); });
        launch(c, "margins_v", [&] { hd_launch_margins(c->d, 0, 2, N, s); });
        launch(c, "smooth_v_hmoi", [&] { hd_launch_normalize_post(c->d, 0, 1, 1, c->cand_cnt, c->cand, c->dryflag, s); });
        swap_v(c);
        launch(c, "wetting_rows", [&] { hd_launch_wetting(c->d, 1, c->cand_cnt, c->cand, c->dryflag, s); });
    } else {
        sidx = 2; eidx = N;
        if (channel && o.kenhhepng == 1) { sidx = 3; eidx = N; }
        if (c->solver == 2) {
            launch(c, "adi_coef_vz", [&] { hd_launch_adi_coef(c->d, 0, sidx, eidx, s); });
            launch(c, "adi_solve_vz", [&] { hd_launch_adi_solve(c->d, 0, sidx, eidx, s); });
        } else if (c->solver == 1) {
            launch(c, "adi_line_vz", [&] { hd_launch_adi_line(c->d, 0, sidx, eidx, s); });
        } else {
            launch(c, "adi_run_vz", [&] { hd_launch_adi_run(c->d, 0, sidx, eidx, c->run_blocks_v, s); });
        }
        launch(c, "smooth_v", [&] { hd_launch_normalize_post(c->d, 0, 0, 0, c->cand_cnt, c->cand, c->dryflag, s); });
        swap_v(c);
        launch(c, "explicit_u", [&] { hd_launch_explicit(c->d, 1, 2, M, s); });
        launch(c, "margins_u", [&] { hd_launch_margins(c->d, 1, 2, M, s); });
        launch(c, "smooth_u_hmoi", [&] { hd_launch_normalize_post(c->d, 1, 1, 0, c->cand_cnt, c->cand, c->dryflag, s); });
        swap_u(c);
        launch(c, "wetting_cols", [&] { hd_launch_wetting(c->d, 0, c->cand_cnt, c->cand, c->dryflag, s); });
    }
    launch(c, "uvz_dry_htd", [&] { hd_launch_uvz_dry(c->d, o.kenhhepd, o.kenhhepng, c->dryflag, s); });
    launch(c, "segments", [&] { hd_launch_segments(c->d, count_wet, s); });
    launch(c, "fs_edges", [&] { hd_launch_fs_edges(c->d, s); });
    if (sediment) {
        launch(c, "vth_hesok", [&] { hd_launch_vth_hesok(c->d, s); });
        if (c->solver_v1) {
            launch(c, which == 0 ? "fs_coef_j" : "fs_coef_i",
                   [&] { hd_launch_fs_coef(c->d, which, c->tdev, o.bed_change_start, o.ketdinh, s); });
            launch(c, which == 0 ? "fs_solve_j" : "fs_solve_i", [&] { hd_launch_fs_solve(c->d, which, s); });
        } else {
            launch(c, which == 0 ? "fs_line_j" : "fs_line_i",
                   [&] { hd_launch_fs_line(c->d, which, c->tdev, o.bed_change_start, o.ketdinh, s); });
        }
        launch(c, "update_fs", [&] { hd_launch_update_fs(c->d, s); });
        if (which == 1) { sidx = 3; eidx = N - 1; }
    }
    if (sidx_out) *sidx_out = sidx;
    if (eidx_out) *eidx_out = eidx;
}

static void enqueue_bed(hd_ctx* c, const hd_step_opts& o, int sidx, int eidx)
{
    launch(c, "calc_qb", [&] { hd_launch_calc_qb(c->d, o.qb_mode, c->stream); });
    launch(c, "bedload", [&] { hd_launch_bedload(c->d, o.ketdinh, sidx, eidx, c->stream); });
}

static void enqueue_step(hd_ctx* c, const hd_step_opts& o, const GraphKey& k)
{
    int sidx = 0, eidx = 0;
    enqueue_half(c, 0, o, k.sed0, 0, nullptr, nullptr);
    enqueue_half(c, 1, o, k.sed1, k.count, &sidx, &eidx);
    if (k.bed) enqueue_bed(c, o, sidx, eidx);
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
    return k;
}

int hd_step(hd_ctx* c, double t, const hd_step_opts* o, double* t_out)
{
    if (!c || !o) return HD_EINVAL;
    if (!c->initialised) return fail(c, HD_ESTATE, "hd_step before hd_init");
    HIPCHK(c, hipSetDevice(c->device));
    double t1 = t + c->P.dT * 0.5, t2 = t1 + c->P.dT * 0.5;
    int rc;
    if ((rc = check_series(c, t1, *o)) || (rc = check_series(c, t2, *o))) return rc;
    hipLaunchKernelGGL(k_set_t, dim3(1), dim3(1), 0, c->stream, c->tdev, t);
    GraphKey k = make_key(c, t, *o, 0);
    enqueue_step(c, *o, k);
    HIPCHK(c, hipGetLastError());
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
    hipLaunchKernelGGL(k_set_t, dim3(1), dim3(1), 0, c->stream, c->tdev, t - c->P.dT * 0.5);
    enqueue_half(c, which, *o, t >= o->sediment_start, 0, nullptr, nullptr);
    HIPCHK(c, hipGetLastError());
    return HD_OK;
}

int hd_bed_change(hd_ctx* c, double t, const hd_step_opts* o)
{
    if (!c || !o) return HD_EINVAL;
    (void)t;
    int sidx = 2, eidx = c->N;
    if ((o->kenhhepd || o->kenhhepng) && o->kenhhepng == 1) sidx = 3;
    if (t >= o->sediment_start) { sidx = 3; eidx = c->N - 1; }
    enqueue_bed(c, *o, sidx, eidx);
    HIPCHK(c, hipGetLastError());
    return HD_OK;
}

static int run_impl(hd_ctx* c, double t, int nsteps, const hd_step_opts* o, double* t_out, int count)
{
    if (!c || !o || nsteps < 0) return HD_EINVAL;
    if (!c->initialised) return fail(c, HD_ESTATE, "hd_run before hd_init");
    HIPCHK(c, hipSetDevice(c->device));
    hipLaunchKernelGGL(k_set_t, dim3(1), dim3(1), 0, c->stream, c->tdev, t);
    for (int n = 0; n < nsteps; ++n) {
        double t1 = t + c->P.dT * 0.5, t2 = t1 + c->P.dT * 0.5;
        int rc;
        if ((rc = check_series(c, t1, *o)) || (rc = check_series(c, t2, *o))) return rc;
        GraphKey k = make_key(c, t, *o, count);
        if (c->profile) {
            enqueue_step(c, *o, k);
            HIPCHK(c, hipGetLastError());
            t = t2;
            continue;
        }
        auto it = c->graphs.find(k);
        if (it == c->graphs.end()) {
            HdDev saved = c->d;
            hipGraph_t g;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            enqueue_step(c, *o, k);
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (e != hipSuccess) return fail(c, HD_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
            hipGraphExec_t ge;
            e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) return fail(c, HD_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
            /* a full step swaps each smoothing pair twice: pointers are back */
            if (std::memcmp(&saved, &c->d, sizeof(HdDev)) != 0) {
                hipGraphExecDestroy(ge);
                return fail(c, HD_EHIP, "step changed the buffer assignment");
            }
            it = c->graphs.emplace(k, ge).first;
        }
        HIPCHK(c, hipGraphLaunch(it->second, c->stream));
        t = t2;
    }
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
    HdDev one = c->d;
    one.wet_count = c->d.wet_count + 1;   /* leave the run accumulator alone */
    HIPCHK(c, hipMemcpyAsync(one.wet_count, &z, 8, hipMemcpyHostToDevice, c->stream));
    hd_launch_wet_count(one, c->stream);
    HIPCHK(c, hipMemcpyAsync(out, one.wet_count, 8, hipMemcpyDeviceToHost, c->stream));
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

int hd_profile(hd_ctx* c, int enable)
{
    if (!c) return HD_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto& r : c->pending) { c->pool.push_back(r.a); c->pool.push_back(r.b); }
    c->pending.clear();
    c->acc.clear();
    c->order.clear();
    c->profile = enable != 0;
    return HD_OK;
}

int hd_kernel_times(hd_ctx* c, int n, const char** names, double* total_ms, int* launches)
{
    if (!c) return HD_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto& r : c->pending) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, r.a, r.b));
        auto it = c->acc.find(r.name);
        if (it == c->acc.end()) { c->order.push_back(r.name); it = c->acc.emplace(r.name, std::make_pair(0.0, 0)).first; }
        it->second.first += ms;
        it->second.second += 1;
        c->pool.push_back(r.a);
        c->pool.push_back(r.b);
    }
    c->pending.clear();
    int k = 0;
    for (const std::string& nm : c->order) {
        if (k >= n) break;
        auto& a = c->acc[nm];
        if (names) names[k] = c->acc.find(nm)->first.c_str();
        if (total_ms) total_ms[k] = a.first;
        if (launches) launches[k] = a.second;
        ++k;
    }
    return k;
}

int hd_thomas_batch(int device, int nsys, const int* off, const int* sn, const double* A, const double* B,
                    const double* C, const double* D, double* x, int nrows)
{
    if (nsys < 0 || nrows < 0 || !off || !sn || !A || !B || !C || !D || !x) return fail(nullptr, HD_EINVAL, "bad args");
    for (int s = 0; s < nsys; ++s)
        if (off[s] < 0 || sn[s] >= 0 && off[s] + sn[s] >= nrows) return fail(nullptr, HD_EINVAL, "system out of range");
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
        hd_thomas_batch_device(nsys, doff, dsn, dA, dB, dC, dD, dx, dAp, dBp, 0);
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = fail(nullptr, HD_EHIP, hipGetErrorString(e));
        else hipMemcpy(x, dx, rb, hipMemcpyDeviceToHost);
    }
    for (void* p : bufs) hipFree(p);
    return rc;
}

} // extern "C"
