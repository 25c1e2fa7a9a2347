// C ABI of liblbm (include/lbm.h): handle, device memory, and the mapping of
// the reference's subroutines onto kernel launches.
//
// Macroscopic fields.  The reference keeps rho, ux, uy, uz as module arrays
// that collision_MRT reads (Channel-Flow/collision.f90:56-59) and macrovar,
// rhoupdat, avedensity, initvel/initpop write.  In the main loop they always
// hold macrovar of the current f, so the step kernel recomputes them from the
// populations it already gathers and the arrays are never touched.  The
// handle tracks what the arrays logically hold:
//   FROM_CUR   macrovar(F_cur) with rho - rho_off  (nothing stored)
//   FROM_PREV  macrovar(F_prev) with rho - rho_off (after a collision that
//              was not followed by macrovar; F_prev is still in the other
//              lattice)
//   ARRAYS     the device arrays hold the values
// and materialises the arrays only when a call needs stored values.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "kernels.h"
#include "lbm.h"

using namespace d3q19;

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(call)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return fail(LBM_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

enum MacroState { FROM_CUR = 0, FROM_PREV = 1, ARRAYS = 2 };

constexpr uint32_t RED_BLOCKS = 2048;   // partial sums per reduction
constexpr uint32_t PLANE_CHUNKS = 512;  // row chunks for plane sums
constexpr uint64_t COPY_NODES = 1u << 20; // nodes per upload/download chunk

} // namespace

struct lbm_channel {
    lbm_config cfg;
    Geom g;
    int prec;
    size_t elem; // bytes per stored population
    void *lat[2];
    int cur;       // lattice holding P of the current state
    bool have_prev;
    bool fresh;    // no step or field assignment yet
    double *mac[4]; // rho, ux, uy, uz (allocated on first use)
    MacroState ms;
    double rho_off;
    double force[3];
    hipStream_t stream;
    double *partials;   // RED_BLOCKS doubles, then PLANE_CHUNKS*11*lx doubles
    double *red_out;    // 1 double + 11*lx doubles of final sums
    unsigned long long *errbits;
    double *stat_acc;   // 11*lx running sums
    long long stat_n;
    double *stage;      // COPY_NODES*19 doubles
    hipEvent_t ev[2];
};

namespace {

// Final sums of per-block partials: out[j] = sum_i part[i*stride + j] for
// j < width, in increasing i (one thread per j) -- deterministic.
__global__ void k_sum_columns(const double *__restrict__ part, uint32_t rows, uint32_t width, double *out,
                              double *acc)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= width)
        return;
    double s = 0.0;
    for (uint32_t i = 0; i < rows; i++)
        s += part[(uint64_t)i * width + j];
    out[j] = s;
    if (acc)
        acc[j] += s;
}

__global__ void k_sub_scalar(double *__restrict__ a, uint64_t n, double v)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        a[i] = a[i] - v;
}

__global__ void k_fill(double *__restrict__ a, uint64_t n, double v)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        a[i] = v;
}

int dispatch_err(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        return fail(LBM_ERR_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
    return LBM_OK;
}

#define PREC_CALL(h, fn, ...)                                                                   \
    ((h)->prec == LBM_FP64   ? Launch<0>::fn(__VA_ARGS__)                                       \
     : (h)->prec == LBM_FP32 ? Launch<1>::fn(__VA_ARGS__)                                       \
                             : Launch<2>::fn(__VA_ARGS__))

RatesD rates_of(const lbm_channel *h)
{
    const lbm_config &c = h->cfg;
    RatesD r;
    r.s1 = c.s1;
    r.s2 = c.s2;
    r.s4 = c.s4;
    r.s9 = c.s9;
    r.s10 = c.s10;
    r.s13 = c.s13;
    r.s16 = c.s16;
    r.omegepsl = c.omegepsl;
    r.omegepslj = c.omegepslj;
    r.omegxx = c.omegxx;
    r.fx = h->force[0];
    r.fy = h->force[1];
    r.fz = h->force[2];
    return r;
}

int ensure_fields(lbm_channel *h)
{
    if (h->mac[0])
        return LBM_OK;
    for (int k = 0; k < 4; k++) {
        hipError_t e = hipMalloc((void **)&h->mac[k], h->g.N * sizeof(double));
        if (e != hipSuccess) {
            for (int j = 0; j < k; j++) {
                (void)hipFree(h->mac[j]);
                h->mac[j] = nullptr;
            }
            return fail(LBM_ERR_NOMEM, "hipMalloc of the macroscopic fields (%zu bytes each) failed: %s",
                        (size_t)(h->g.N * sizeof(double)), hipGetErrorString(e));
        }
        HIP_TRY(hipMemsetAsync(h->mac[k], 0, h->g.N * sizeof(double), h->stream));
    }
    return LBM_OK;
}

// macrovar of lattice `which` (minus the pending rho offset) into the arrays.
int compute_fields(lbm_channel *h, int which)
{
    int rc = ensure_fields(h);
    if (rc)
        return rc;
    MacroArgs a;
    a.g = h->g;
    a.L = h->lat[which];
    a.rho = h->mac[0];
    a.ux = h->mac[1];
    a.uy = h->mac[2];
    a.uz = h->mac[3];
    a.fx = h->force[0];
    a.fy = h->force[1];
    a.fz = h->force[2];
    a.rho_off = h->rho_off;
    a.errbits = nullptr;
    return dispatch_err(PREC_CALL(h, macro, a, 0, h->stream), "macrovar");
}

// Make the arrays hold what the reference's arrays hold now.
int materialize(lbm_channel *h)
{
    if (h->ms == ARRAYS)
        return LBM_OK;
    const int which = (h->ms == FROM_CUR) ? h->cur : 1 - h->cur;
    if (h->ms == FROM_PREV && !h->have_prev)
        return fail(LBM_ERR_STATE, "internal: previous lattice not available");
    int rc = compute_fields(h, which);
    if (rc)
        return rc;
    h->ms = ARRAYS;
    h->rho_off = 0.0;
    return LBM_OK;
}

int launch_step(lbm_channel *h)
{
    StepArgs a;
    a.g = h->g;
    a.src = h->lat[h->cur];
    a.dst = h->lat[1 - h->cur];
    a.rho = h->mac[0];
    a.ux = h->mac[1];
    a.uy = h->mac[2];
    a.uz = h->mac[3];
    a.rho_off = h->rho_off;
    a.r = rates_of(h);
    int mode;
    if (h->ms == FROM_PREV) {
        int rc = materialize(h);
        if (rc)
            return rc;
        a.rho = h->mac[0];
        a.ux = h->mac[1];
        a.uy = h->mac[2];
        a.uz = h->mac[3];
        a.rho_off = 0.0;
    }
    if (h->ms == ARRAYS)
        mode = 2;
    else
        mode = (h->rho_off != 0.0) ? 1 : 0;
    int rc = dispatch_err(PREC_CALL(h, step, a, mode, h->stream), "collision");
    if (rc)
        return rc;
    h->cur = 1 - h->cur;
    h->have_prev = true;
    h->fresh = false;
    if (h->ms == FROM_CUR)
        h->ms = FROM_PREV;
    return LBM_OK;
}

int reduce_rho(lbm_channel *h, int src, double *sum)
{
    SumArgs a;
    a.g = h->g;
    a.L = h->lat[h->cur];
    a.rho = h->mac[0];
    a.partials = h->partials;
    uint32_t grid = (uint32_t)std::min<uint64_t>(RED_BLOCKS, (h->g.N + 255) / 256);
    int rc = dispatch_err(PREC_CALL(h, rho_partials, a, src, grid, h->stream), "rho reduction");
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_sum_columns, dim3(1), dim3(64), 0, h->stream, h->partials, grid, 1u, h->red_out,
                       (double *)nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(sum, h->red_out, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

// Plane sums of the current fields into red_out[0 .. 11*lx) (and the running
// accumulator when acc is set).
int plane_sums(lbm_channel *h, bool acc)
{
    PlaneArgs a;
    a.g = h->g;
    a.L = h->lat[h->ms == FROM_PREV ? 1 - h->cur : h->cur];
    a.rho = h->mac[0];
    a.ux = h->mac[1];
    a.uy = h->mac[2];
    a.uz = h->mac[3];
    a.fx = h->force[0];
    a.fy = h->force[1];
    a.fz = h->force[2];
    a.rho_off = h->rho_off;
    a.partials = h->partials + RED_BLOCKS;
    const uint32_t rows = h->g.ly * h->g.lz;
    const uint32_t chunks = std::min<uint32_t>(PLANE_CHUNKS, rows);
    const int src = (h->ms == ARRAYS) ? 0 : 1;
    int rc = dispatch_err(PREC_CALL(h, plane_partials, a, src, chunks, h->stream), "plane sums");
    if (rc)
        return rc;
    const uint32_t width = 11 * h->g.lx;
    hipLaunchKernelGGL(k_sum_columns, dim3((width + 255) / 256), dim3(256), 0, h->stream, a.partials, chunks, width,
                       h->red_out + 1, acc ? h->stat_acc : (double *)nullptr);
    HIP_TRY(hipGetLastError());
    return LBM_OK;
}

// Columns of saveload.f90:1430-1492 from plane sums s[q*lx + i] over nf
// samples per plane.
void stat_columns(const lbm_channel *h, const double *s, double nf, double ustar, double ystar, double *out)
{
    const uint32_t lx = h->g.lx;
    const double us2 = ustar * ustar, us4 = us2 * us2;
    const double volf = 1.0; // every node is fluid
    for (uint32_t i = 0; i < lx; i++) {
        const double *c = s + i;
        double vx = c[0 * lx] / nf / ustar, vy = c[1 * lx] / nf / ustar, vz = c[2 * lx] / nf / ustar;
        double sxz = c[3 * lx] / nf / us2 - vx * vz;
        double sxy = c[4 * lx] / nf / us2 - vx * vy;
        double syz = c[5 * lx] / nf / us2 - vy * vz;
        double xx = c[6 * lx] / nf / us2 - vx * vx;
        double yy = c[7 * lx] / nf / us2 - vy * vy;
        double zz = c[8 * lx] / nf / us2 - vz * vz;
        double pr = c[9 * lx] / nf / us2;
        double p2 = c[10 * lx] / nf / us4 - pr * pr;
        double yplus = ((double)(i + 1) - 0.5) / ystar;
        if (i + 1 > lx / 2)
            yplus = ((double)(lx - (i + 1)) + 0.5) / ystar;
        double *o = out + 15 * (size_t)i;
        o[0] = (double)(i + 1) - 0.5;
        o[1] = yplus;
        o[2] = vx;
        o[3] = vy;
        o[4] = vz;
        o[5] = xx;
        o[6] = yy;
        o[7] = zz;
        o[8] = sxz;
        o[9] = sxy;
        o[10] = syz;
        o[11] = pr;
        o[12] = p2;
        o[13] = volf;
        o[14] = 1.0 - volf;
    }
}

bool same(double a, double b) { return a == b; }

} // namespace

extern "C" {

const char *lbm_last_error(void) { return g_err; }

int lbm_create(const lbm_config *cfg, lbm_channel **out)
{
    g_err[0] = 0;
    if (!cfg || !out)
        return fail(LBM_ERR_ARG, "lbm_create: null argument");
    *out = nullptr;
    const lbm_config &c = *cfg;
    if (c.lx < 2 || c.ly < 1 || c.lz < 1)
        return fail(LBM_ERR_ARG, "lbm_create: grid %d x %d x %d: need lx >= 2, ly >= 1, lz >= 1", c.lx, c.ly, c.lz);
    const uint64_t N = (uint64_t)c.lx * c.ly * c.lz;
    if (N >= (1ull << 31))
        return fail(LBM_ERR_ARG, "lbm_create: %llu nodes exceed the 2^31 per-device limit of 32-bit node indices",
                    (unsigned long long)N);
    if (c.precision < LBM_FP64 || c.precision > LBM_FP16)
        return fail(LBM_ERR_ARG, "lbm_create: unknown precision %d", c.precision);
    if (c.reserved != 0)
        return fail(LBM_ERR_ARG, "lbm_create: reserved field must be 0");
    // structural constants are compiled into the kernels (para.f90:143-174)
    struct {
        const char *name;
        double got, want;
    } chk[] = {{"coef1", c.coef1, -(2.0 / 3.0)}, {"coef2", c.coef2, -11.0},        {"coef3", c.coef3, 8.0},
               {"coef4", c.coef4, -4.0},         {"coef5", c.coef5, 2.0},          {"coef3i", c.coef3i, 1.0 / 8.0},
               {"coef4i", c.coef4i, 1.0 / -4.0}, {"val8", c.val8, 12.0},           {"val1i", c.val1i, 1.0 / 19.0},
               {"val2i", c.val2i, 1.0 / 2394.0}, {"val3i", c.val3i, 1.0 / 252.0},  {"val4i", c.val4i, 1.0 / 10.0},
               {"val5i", c.val5i, 1.0 / 40.0},   {"val6i", c.val6i, 1.0 / 36.0},   {"val7i", c.val7i, 1.0 / 72.0},
               {"val8i", c.val8i, 1.0 / 12.0},   {"val9i", c.val9i, 1.0 / 24.0},   {"ww0", c.ww0, 1.0 / 3.0},
               {"ww1", c.ww1, 1.0 / 18.0},       {"ww2", c.ww2, 1.0 / 36.0}};
    for (auto &k : chk)
        if (!same(k.got, k.want))
            return fail(LBM_ERR_ARG, "lbm_create: %s = %.17g differs from Channel-Flow/para.f90's %.17g", k.name,
                        k.got, k.want);

    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (c.device < 0 || c.device >= ndev)
        return fail(LBM_ERR_ARG, "lbm_create: device %d not present (%d devices)", c.device, ndev);
    HIP_TRY(hipSetDevice(c.device));

    lbm_channel *h = new lbm_channel();
    memset(h, 0, sizeof(*h));
    h->cfg = c;
    h->prec = c.precision;
    h->elem = c.precision == LBM_FP64 ? 8 : (c.precision == LBM_FP32 ? 4 : 2);
    h->g.lx = c.lx;
    h->g.ly = c.ly;
    h->g.lz = c.lz;
    h->g.plane = (uint32_t)c.lx * c.ly;
    h->g.N = N;
    h->g.dlx = make_fastdiv(c.lx);
    h->g.dly = make_fastdiv(c.ly);
    h->g.remap = 0;
    h->ms = FROM_CUR;
    h->fresh = true;

    auto cleanup = [&](int code) {
        lbm_destroy(h);
        return code;
    };
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess)
        return cleanup(fail(LBM_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e)));
    const size_t lat_bytes = (size_t)N * 19 * h->elem;
    for (int k = 0; k < 2; k++) {
        e = hipMalloc(&h->lat[k], lat_bytes);
        if (e != hipSuccess)
            return cleanup(fail(LBM_ERR_NOMEM, "hipMalloc of lattice %d (%zu bytes) failed: %s", k, lat_bytes,
                                hipGetErrorString(e)));
        e = hipMemsetAsync(h->lat[k], 0, lat_bytes, h->stream);
        if (e != hipSuccess)
            return cleanup(fail(LBM_ERR_HIP, "hipMemsetAsync failed: %s", hipGetErrorString(e)));
    }
    const size_t part = RED_BLOCKS + (size_t)PLANE_CHUNKS * 11 * c.lx;
    if (hipMalloc((void **)&h->partials, part * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&h->red_out, (1 + 11 * (size_t)c.lx) * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&h->stat_acc, 11 * (size_t)c.lx * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&h->errbits, sizeof(unsigned long long)) != hipSuccess)
        return cleanup(fail(LBM_ERR_NOMEM, "hipMalloc of reduction scratch failed"));
    e = hipMemsetAsync(h->stat_acc, 0, 11 * (size_t)c.lx * sizeof(double), h->stream);
    if (e == hipSuccess)
        e = hipEventCreate(&h->ev[0]);
    if (e == hipSuccess)
        e = hipEventCreate(&h->ev[1]);
    if (e == hipSuccess)
        e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess)
        return cleanup(fail(LBM_ERR_HIP, "lbm_create setup failed: %s", hipGetErrorString(e)));
    *out = h;
    return LBM_OK;
}

int lbm_destroy(lbm_channel *h)
{
    if (!h)
        return LBM_OK;
    (void)hipSetDevice(h->cfg.device);
    if (h->stream)
        (void)hipStreamSynchronize(h->stream);
    for (int k = 0; k < 2; k++)
        if (h->lat[k])
            (void)hipFree(h->lat[k]);
    for (int k = 0; k < 4; k++)
        if (h->mac[k])
            (void)hipFree(h->mac[k]);
    if (h->partials)
        (void)hipFree(h->partials);
    if (h->red_out)
        (void)hipFree(h->red_out);
    if (h->stat_acc)
        (void)hipFree(h->stat_acc);
    if (h->errbits)
        (void)hipFree(h->errbits);
    if (h->stage)
        (void)hipFree(h->stage);
    if (h->ev[0])
        (void)hipEventDestroy(h->ev[0]);
    if (h->ev[1])
        (void)hipEventDestroy(h->ev[1]);
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
    return LBM_OK;
}

#define CHECK_H(h)                                                       \
    do {                                                                 \
        if (!(h))                                                        \
            return fail(LBM_ERR_ARG, "%s: null handle", __func__);       \
        g_err[0] = 0;                                                    \
        HIP_TRY(hipSetDevice((h)->cfg.device));                          \
    } while (0)

static int ensure_stage(lbm_channel *h)
{
    if (h->stage)
        return LBM_OK;
    hipError_t e = hipMalloc((void **)&h->stage, COPY_NODES * 19 * sizeof(double));
    if (e != hipSuccess)
        return fail(LBM_ERR_NOMEM, "hipMalloc of the copy staging buffer failed: %s", hipGetErrorString(e));
    return LBM_OK;
}

int lbm_upload_f(lbm_channel *h, const double *f)
{
    CHECK_H(h);
    if (!f)
        return fail(LBM_ERR_ARG, "lbm_upload_f: null f");
    int rc = ensure_stage(h);
    if (rc)
        return rc;
    // the reference's arrays keep their values when f is replaced
    if (!h->fresh && h->ms != ARRAYS) {
        rc = materialize(h);
        if (rc)
            return rc;
    }
    for (uint64_t n0 = 0; n0 < h->g.N; n0 += COPY_NODES) {
        const uint64_t cnt = std::min<uint64_t>(COPY_NODES, h->g.N - n0);
        HIP_TRY(hipMemcpyAsync(h->stage, f + n0 * 19, cnt * 19 * sizeof(double), hipMemcpyHostToDevice, h->stream));
        CopyArgs a;
        a.g = h->g;
        a.L = h->lat[h->cur];
        a.aos = h->stage;
        a.n0 = n0;
        a.count = cnt;
        rc = dispatch_err(PREC_CALL(h, scatter, a, h->stream), "upload scatter");
        if (rc)
            return rc;
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->have_prev = false;
    return LBM_OK;
}

int lbm_download_f(lbm_channel *h, double *f)
{
    CHECK_H(h);
    if (!f)
        return fail(LBM_ERR_ARG, "lbm_download_f: null f");
    int rc = ensure_stage(h);
    if (rc)
        return rc;
    for (uint64_t n0 = 0; n0 < h->g.N; n0 += COPY_NODES) {
        const uint64_t cnt = std::min<uint64_t>(COPY_NODES, h->g.N - n0);
        CopyArgs a;
        a.g = h->g;
        a.L = h->lat[h->cur];
        a.aos = h->stage;
        a.n0 = n0;
        a.count = cnt;
        rc = dispatch_err(PREC_CALL(h, gather, a, h->stream), "download gather");
        if (rc)
            return rc;
        HIP_TRY(hipMemcpyAsync(f + n0 * 19, h->stage, cnt * 19 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    return LBM_OK;
}

int lbm_set_force(lbm_channel *h, double fx, double fy, double fz)
{
    CHECK_H(h);
    // the arrays the reference holds were formed with the old force
    if (h->ms != ARRAYS && !h->fresh && (fx != h->force[0] || fy != h->force[1] || fz != h->force[2])) {
        int rc = materialize(h);
        if (rc)
            return rc;
    }
    h->force[0] = fx;
    h->force[1] = fy;
    h->force[2] = fz;
    return LBM_OK;
}

int lbm_set_macros(lbm_channel *h, const double *rho, const double *ux, const double *uy, const double *uz)
{
    CHECK_H(h);
    int rc = materialize(h);
    if (rc)
        return rc;
    const double *src[4] = {rho, ux, uy, uz};
    for (int k = 0; k < 4; k++)
        if (src[k])
            HIP_TRY(hipMemcpyAsync(h->mac[k], src[k], h->g.N * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->fresh = false;
    return LBM_OK;
}

int lbm_initpop(lbm_channel *h)
{
    CHECK_H(h);
    int rc = materialize(h);
    if (rc)
        return rc;
    InitArgs a;
    a.g = h->g;
    a.L = h->lat[h->cur];
    a.rho = h->mac[0];
    a.ux = h->mac[1];
    a.uy = h->mac[2];
    a.uz = h->mac[3];
    rc = dispatch_err(PREC_CALL(h, initpop, a, h->stream), "initpop");
    if (rc)
        return rc;
    h->have_prev = false;
    h->fresh = false;
    return LBM_OK;
}

int lbm_collision_mrt(lbm_channel *h)
{
    CHECK_H(h);
    return launch_step(h);
}

int lbm_macrovar(lbm_channel *h, double *rho, double *ux, double *uy, double *uz)
{
    CHECK_H(h);
    h->ms = FROM_CUR;
    h->rho_off = 0.0;
    h->fresh = false;
    double *dst[4] = {rho, ux, uy, uz};
    if (!rho && !ux && !uy && !uz)
        return LBM_OK;
    int rc = compute_fields(h, h->cur);
    if (rc)
        return rc;
    for (int k = 0; k < 4; k++)
        if (dst[k])
            HIP_TRY(hipMemcpyAsync(dst[k], h->mac[k], h->g.N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_rhoupdat(lbm_channel *h, double *rhoerrmax)
{
    CHECK_H(h);
    // u keeps what the arrays hold; rho becomes sum f of the current lattice
    int rc = materialize(h);
    if (rc)
        return rc;
    HIP_TRY(hipMemsetAsync(h->errbits, 0, sizeof(unsigned long long), h->stream));
    MacroArgs a;
    a.g = h->g;
    a.L = h->lat[h->cur];
    a.rho = h->mac[0];
    a.ux = h->mac[1];
    a.uy = h->mac[2];
    a.uz = h->mac[3];
    a.fx = h->force[0];
    a.fy = h->force[1];
    a.fz = h->force[2];
    a.rho_off = 0.0;
    a.errbits = h->errbits;
    rc = dispatch_err(PREC_CALL(h, macro, a, 1, h->stream), "rhoupdat");
    if (rc)
        return rc;
    h->fresh = false;
    if (rhoerrmax) {
        unsigned long long bits = 0;
        HIP_TRY(hipMemcpyAsync(&bits, h->errbits, sizeof(bits), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        memcpy(rhoerrmax, &bits, sizeof(double));
    }
    return LBM_OK;
}

int lbm_avedensity(lbm_channel *h, double *rhomean)
{
    CHECK_H(h);
    double sum = 0.0, mean;
    int rc;
    if (h->ms == FROM_CUR && h->rho_off == 0.0) {
        // rho is virtual: reduce macrovar(F) directly and keep the correction
        // as an offset applied by the next collision
        rc = reduce_rho(h, 1, &sum);
        if (rc)
            return rc;
        mean = sum / (double)h->g.N;
        h->rho_off = mean;
    } else {
        rc = materialize(h);
        if (rc)
            return rc;
        rc = reduce_rho(h, 0, &sum);
        if (rc)
            return rc;
        mean = sum / (double)h->g.N;
        hipLaunchKernelGGL(k_sub_scalar, dim3((unsigned)((h->g.N + 255) / 256)), dim3(256), 0, h->stream, h->mac[0],
                           h->g.N, mean);
        HIP_TRY(hipGetLastError());
    }
    h->fresh = false;
    if (rhomean)
        *rhomean = mean;
    return LBM_OK;
}

int lbm_run(lbm_channel *h, int nsteps)
{
    CHECK_H(h);
    if (nsteps < 0)
        return fail(LBM_ERR_ARG, "lbm_run: nsteps %d < 0", nsteps);
    for (int s = 0; s < nsteps; s++) {
        int rc = launch_step(h);
        if (rc)
            return rc;
        h->ms = FROM_CUR; // macrovar
        h->rho_off = 0.0;
    }
    return LBM_OK;
}

int lbm_mass_sum(lbm_channel *h, double *mass)
{
    CHECK_H(h);
    if (!mass)
        return fail(LBM_ERR_ARG, "lbm_mass_sum: null output");
    SumArgs a;
    a.g = h->g;
    a.L = h->lat[h->cur];
    a.rho = nullptr;
    a.partials = h->partials;
    uint32_t grid = (uint32_t)std::min<uint64_t>(RED_BLOCKS, (h->g.N + 255) / 256);
    int rc = dispatch_err(PREC_CALL(h, mass_partials, a, grid, h->stream), "mass reduction");
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_sum_columns, dim3(1), dim3(64), 0, h->stream, h->partials, grid, 1u, h->red_out,
                       (double *)nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(mass, h->red_out, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_statistc2(lbm_channel *h, double ustar, double ystar, double *out)
{
    CHECK_H(h);
    if (!out)
        return fail(LBM_ERR_ARG, "lbm_statistc2: null output");
    int rc = plane_sums(h, false);
    if (rc)
        return rc;
    std::vector<double> s(11 * (size_t)h->g.lx);
    HIP_TRY(hipMemcpyAsync(s.data(), h->red_out + 1, s.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    stat_columns(h, s.data(), (double)h->g.ly * (double)h->g.lz, ustar, ystar, out);
    return LBM_OK;
}

int lbm_stats_accumulate(lbm_channel *h)
{
    CHECK_H(h);
    int rc = plane_sums(h, true);
    if (rc)
        return rc;
    h->stat_n++;
    return LBM_OK;
}

int lbm_stats_read(lbm_channel *h, double ustar, double ystar, double *out, long long *nsamples)
{
    CHECK_H(h);
    if (!out)
        return fail(LBM_ERR_ARG, "lbm_stats_read: null output");
    if (nsamples)
        *nsamples = h->stat_n;
    if (h->stat_n == 0)
        return fail(LBM_ERR_STATE, "lbm_stats_read: no samples accumulated");
    std::vector<double> s(11 * (size_t)h->g.lx);
    HIP_TRY(hipMemcpyAsync(s.data(), h->stat_acc, s.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    stat_columns(h, s.data(), (double)h->g.ly * (double)h->g.lz * (double)h->stat_n, ustar, ystar, out);
    return LBM_OK;
}

int lbm_stats_reset(lbm_channel *h)
{
    CHECK_H(h);
    const uint64_t n = 11 * (uint64_t)h->g.lx;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->stat_acc, n, 0.0);
    HIP_TRY(hipGetLastError());
    h->stat_n = 0;
    return LBM_OK;
}

int lbm_sync(lbm_channel *h)
{
    CHECK_H(h);
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_time_steps(lbm_channel *h, int nsteps, double *avg_ms, double *total_ms)
{
    CHECK_H(h);
    if (nsteps <= 0)
        return fail(LBM_ERR_ARG, "lbm_time_steps: nsteps must be positive");
    std::vector<hipEvent_t> ev(2 * (size_t)nsteps);
    for (auto &e : ev)
        HIP_TRY(hipEventCreate(&e));
    int rc = LBM_OK;
    for (int s = 0; s < nsteps && rc == LBM_OK; s++) {
        HIP_TRY(hipEventRecord(ev[2 * s], h->stream));
        rc = launch_step(h);
        h->ms = FROM_CUR;
        h->rho_off = 0.0;
        HIP_TRY(hipEventRecord(ev[2 * s + 1], h->stream));
    }
    if (rc == LBM_OK) {
        HIP_TRY(hipEventSynchronize(ev[2 * nsteps - 1]));
        double sum = 0.0;
        for (int s = 0; s < nsteps; s++) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]));
            sum += ms;
        }
        float tot = 0.f;
        HIP_TRY(hipEventElapsedTime(&tot, ev[0], ev[2 * nsteps - 1]));
        if (avg_ms)
            *avg_ms = sum / nsteps;
        if (total_ms)
            *total_ms = tot;
    }
    for (auto &e : ev)
        (void)hipEventDestroy(e);
    return rc;
}

int lbm_device_bytes(lbm_channel *h, size_t *bytes)
{
    if (!h || !bytes)
        return fail(LBM_ERR_ARG, "lbm_device_bytes: null argument");
    size_t b = 2 * (size_t)h->g.N * 19 * h->elem;
    if (h->mac[0])
        b += 4 * (size_t)h->g.N * sizeof(double);
    if (h->stage)
        b += COPY_NODES * 19 * sizeof(double);
    b += (RED_BLOCKS + (size_t)PLANE_CHUNKS * 11 * h->g.lx + 1 + 22 * (size_t)h->g.lx) * sizeof(double) + 8;
    *bytes = b;
    return LBM_OK;
}

} // extern "C"
