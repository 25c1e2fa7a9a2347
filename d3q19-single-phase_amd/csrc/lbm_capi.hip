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

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "blocked.h"
#include "kernels.h"
#include "lbm.h"
#include "particles.h"
#include "mlp_train.h"

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

// Scratch of lbm_mass_sum_noalloc: the per-block partial sums of the mass
// reduction, 16 KiB per GPU, held until the process ends.  Host code never passes `scratch` itself
// to a kernel: it uses the device address that hipGetSymbolAddress returns
// for the current device (scratch_ptr), kept per device.
constexpr uint32_t NOALLOC_PARTIALS = RED_BLOCKS;
constexpr int MAX_DEVICES = 64;
double *g_scratch_ptr[MAX_DEVICES];
std::mutex g_scratch_mu;

} // namespace

static __device__ __align__(256) double scratch[NOALLOC_PARTIALS];

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
    ForceSpec fc;       // the force arrays' contents: FORCING (uniform) or FORCINGP
    ForceSpec fm;       // the force that the fields FROM_CUR / FROM_PREV are formed with
    double *ftab;       // FORCINGP factor tables (allocated on first use)
    hipStream_t stream;
    double *partials;   // 2*RED_BLOCKS doubles, then PLANE_CHUNKS*NPLANE*lx doubles
    double *red_out;    // 2 doubles + NPLANE*lx doubles of final sums
    unsigned long long *errbits;
    double *stat_acc;   // NPLANE*lx running sums
    long long stat_n;
    double *fstat_acc;  // NPLANE*(lx/2) running sums of the box-filtered fields (allocated on first use)
    long long fstat_n;
    double *stage;      // COPY_NODES*19 doubles
    hipEvent_t ev[2];
    struct Part *part;  // rigid particles (nullptr without)
    hipStream_t pstream; // lbm_run_particles: the force sums and beads_move beside the collision
    hipEvent_t pev[2];   // its fork (after the bounce-back) and join (before beads_links)
    float *mlp;         // learned closure parameters on the device (nullptr: plain MRT)
    float mlp_inv_ystar, mlp_yscale;
    int runs;           // step kernel: 1 = vector runs where the shape allows, 0 = one node per thread
    int w32;            // step kernel: 1 = 32-bit pull offsets where N allows (default), 2 = between adjacent rows only
    uint32_t bK;        // two-step blocked sweep: slabs per chunk (0 = single steps only)
    int baxis;          // its sweep axis: 2 = z (slabs lx*ly), 1 = y (slabs lx*lz)
    uint32_t bW;        // rows per window of its slabs (0 = whole slabs)
    int bforced;        // lbm_run_forcingp also pairs steps whose force varies over the nodes
    void *bring, *bhead; // its slab ring and head buffer (allocated on first use)
    bool ave_pending;   // lbm_run_particles: avedensity due once the next bounce-back has run
    // Per-call pairing: collision_MRT calls whose launch is deferred (0-2) so
    // that collision, macrovar, collision, macrovar runs as one blocked pair;
    // pend_mac: a fields-only macrovar followed the first; defer: enabled
    // (LBM_DEFER=0 at lbm_create turns it off)
    int pend;
    bool pend_mac;
    int defer;
};

struct lbm_trainer {
    int device;
    hipStream_t stream;
    long long capacity, count;
    float *feat, *target;   // capacity x 10, capacity
    float *w, *m, *v;       // LBM_MLP_TOTAL parameters, Adam moments
    float *partial, *grad;  // per-wave partial gradients, summed gradient (+ loss)
    uint32_t max_waves;
    int step;
    int target_mode; // LBM_TARGET_LOCAL or LBM_TARGET_PLANE for lbm_trainer_add_filtered
};

// Particle set on the device (see particles.h).
struct Part {
    PartGeom pg;
    PartState ps;
    std::vector<void *> allocs;
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
     : (h)->prec == LBM_FP16 ? Launch<2>::fn(__VA_ARGS__)                                       \
                             : Launch<3>::fn(__VA_ARGS__))

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
    r.fx = h->fc.fx;
    r.fy = h->fc.fy;
    r.fz = h->fc.fz;
    return with_f32(r);
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
    a.f = h->fm;
    a.rho_off = h->rho_off;
    a.errbits = nullptr;
    rc = dispatch_err(PREC_CALL(h, macro, a, 0, h->stream), "macrovar");
    if (rc || !h->part)
        return rc;
    // macrovar's solid branch (collision.f90:420-459), with the current particle state
    const Part *pt = h->part;
    return dispatch_err(part_solid_fields(pt->pg, pt->ps, h->rho_off, h->mac[0], h->mac[1], h->mac[2], h->mac[3],
                                          h->stream),
                        "macrovar inside particles");
}

// macrovar of the current lattice (with the force arrays as they are now)
// becomes what the fields hold.
void fields_from_cur(lbm_channel *h)
{
    h->ms = FROM_CUR;
    h->fm = h->fc;
}

// Do two force specifications give the same force at every node?
bool same_force(const ForceSpec &a, const ForceSpec &b)
{
    return a.tab == b.tab && a.fx == b.fx && a.fy == b.fy && a.fz == b.fz &&
           (a.tab == nullptr || (a.a == b.a && a.b == b.b && a.c == b.c));
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
    a.solid = nullptr;
    if (h->part) {
        mode |= 4;
        a.solid = h->part->ps.solid;
    }
    a.fc = h->fc;
    a.fm = h->fm;
    // a force that varies over the nodes, or fused fields formed with a force
    // other than the collision's
    if (h->fc.tab || (h->ms != ARRAYS && !same_force(h->fm, h->fc))) {
        if (h->part)
            return fail(LBM_ERR_STATE, "collision: particles need a uniform force that macrovar has seen "
                                       "(FORCINGP is not combined with particles)");
        mode |= 16;
    }
    a.mlp = h->mlp;
    a.inv_ystar = h->mlp_inv_ystar;
    a.yscale = h->mlp_yscale;
    a.runs = h->runs;
    a.w32 = h->w32;
    if (h->mlp) {
        if (mode & 3) // the closure runs with fused fields only
            return fail(LBM_ERR_STATE, "collision: the learned closure needs fields from macrovar");
        mode |= 8;
    }
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

const int *solid_of(const lbm_channel *h) { return h->part ? h->part->ps.solid : nullptr; }

// Sum of rho over fluid nodes (SRC 0: field array, SRC 1: macrovar of the
// current lattice) and the number of fluid nodes.
int reduce_rho(lbm_channel *h, int src, double *sum, double *count)
{
    SumArgs a;
    a.g = h->g;
    a.L = h->lat[h->cur];
    a.rho = h->mac[0];
    a.solid = solid_of(h);
    a.partials = h->partials;
    uint32_t grid = (uint32_t)std::min<uint64_t>(RED_BLOCKS, (h->g.N + 255) / 256);
    int rc = dispatch_err(PREC_CALL(h, rho_partials, a, src, grid, h->stream), "rho reduction");
    if (rc)
        return rc;
    // partials: sums in [0, grid), counts in [grid, 2*grid)
    hipLaunchKernelGGL(k_sum_columns, dim3(1), dim3(64), 0, h->stream, h->partials, grid, 1u, h->red_out,
                       (double *)nullptr);
    hipLaunchKernelGGL(k_sum_columns, dim3(1), dim3(64), 0, h->stream, h->partials + grid, grid, 1u, h->red_out + 1,
                       (double *)nullptr);
    HIP_TRY(hipGetLastError());
    double out[2];
    HIP_TRY(hipMemcpyAsync(out, h->red_out, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    *sum = out[0];
    *count = out[1];
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
    a.f = h->fm;
    a.rho_off = h->rho_off;
    a.solid = solid_of(h);
    a.partials = h->partials + 2 * RED_BLOCKS;
    const uint32_t rows = h->g.ly * h->g.lz;
    const uint32_t chunks = std::min<uint32_t>(PLANE_CHUNKS, rows);
    const int src = (h->ms == ARRAYS) ? 0 : 1;
    int rc = dispatch_err(PREC_CALL(h, plane_partials, a, src, chunks, h->stream), "plane sums");
    if (rc)
        return rc;
    const uint32_t width = NPLANE * h->g.lx;
    hipLaunchKernelGGL(k_sum_columns, dim3((width + 255) / 256), dim3(256), 0, h->stream, a.partials, chunks, width,
                       h->red_out + 2, acc ? h->stat_acc : (double *)nullptr);
    HIP_TRY(hipGetLastError());
    return LBM_OK;
}

// Plane sums of the 2x2x2 box-filtered fields of the current populations
// into red_out[2 .. 2 + 12*lx/2) and the filtered running accumulator.
int filtered_plane_sums(lbm_channel *h)
{
    PlaneArgs a;
    a.g = h->g;
    a.L = h->lat[h->ms == FROM_PREV ? 1 - h->cur : h->cur];
    a.rho = a.ux = a.uy = a.uz = nullptr;
    a.f = h->fm;
    a.rho_off = h->rho_off;
    a.solid = nullptr;
    a.partials = h->partials + 2 * RED_BLOCKS;
    const uint32_t rows = (h->g.ly / 2) * (h->g.lz / 2);
    const uint32_t chunks = std::min<uint32_t>(PLANE_CHUNKS, rows);
    int rc = dispatch_err(PREC_CALL(h, filtered_plane_partials, a, chunks, h->stream), "filtered plane sums");
    if (rc)
        return rc;
    const uint32_t width = NPLANE * (h->g.lx / 2);
    hipLaunchKernelGGL(k_sum_columns, dim3((width + 255) / 256), dim3(256), 0, h->stream, a.partials, chunks, width,
                       h->red_out + 2, h->fstat_acc);
    HIP_TRY(hipGetLastError());
    return LBM_OK;
}

// Columns of saveload.f90:1430-1492 from plane sums s[q*lx + i] of a grid
// with lx planes of plane_nodes nodes; the fluid-node count of plane i is
// s[11*lx + i] (per sample), nsamp the number of accumulated samples.
void stat_columns(uint32_t lx, double plane_nodes, const double *s, double nsamp, double ustar, double ystar,
                  double *out)
{
    const double us2 = ustar * ustar, us4 = us2 * us2;
    for (uint32_t i = 0; i < lx; i++) {
        const double *c = s + i;
        const double nf = c[11 * lx] > 0.0 ? c[11 * lx] : 1.0;
        double vx = c[0 * lx] / nf / ustar, vy = c[1 * lx] / nf / ustar, vz = c[2 * lx] / nf / ustar;
        double sxz = c[3 * lx] / nf / us2 - vx * vz;
        double sxy = c[4 * lx] / nf / us2 - vx * vy;
        double syz = c[5 * lx] / nf / us2 - vy * vz;
        double xx = c[6 * lx] / nf / us2 - vx * vx;
        double yy = c[7 * lx] / nf / us2 - vy * vy;
        double zz = c[8 * lx] / nf / us2 - vz * vz;
        double pr = c[9 * lx] / nf / us2;
        double p2 = c[10 * lx] / nf / us4 - pr * pr;
        const double volf = c[11 * lx] / nsamp / plane_nodes;
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
    if (c.precision < LBM_FP64 || c.precision > LBM_FP16C)
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
    if (c.device >= MAX_DEVICES)
        return fail(LBM_ERR_ARG, "lbm_create: device %d beyond the library's %d-device table", c.device, MAX_DEVICES);
    HIP_TRY(hipSetDevice(c.device));
    {
        // device address of this GPU's copy of `scratch`, once per device
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[c.device]) {
            void *p = nullptr;
            HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            if (!p)
                return fail(LBM_ERR_HIP, "lbm_create: hipGetSymbolAddress(scratch) gave a null address");
            g_scratch_ptr[c.device] = (double *)p;
        }
    }

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
    {
        // LBM_STEP_KERNEL=runs selects two nodes per thread (measured slower on
        // the 192x768x384 channel: 64.6 % vs 69.2 % of HBM in FP64)
        const char *k = getenv("LBM_STEP_KERNEL");
        h->runs = (k && strcmp(k, "runs") == 0) ? 1 : 0;
        // LBM_STEP_W32=0: 2-byte storage and the FP32 closure address their
        // walled pulls from two 64-bit row bases even where 32-bit offsets
        // would do; LBM_STEP_W32=adjacent: 2-byte storage takes 32-bit
        // offsets between adjacent rows only, the form lattices above 2^29
        // nodes use (the same values; for comparing the forms)
        const char *w = getenv("LBM_STEP_W32");
        h->w32 = (w && strcmp(w, "0") == 0) ? 0 : (w && strcmp(w, "adjacent") == 0) ? 2 : 1;
        // two-step blocked sweep (blocked.h, block_choice): chunks of K slabs
        // whose ring of K + 2 slabs stays within BLOCK_RING_BYTES, along the
        // axis whose chunk holds more nodes.  Below BLOCK_MIN_SLABS whole
        // slabs measured slower than single steps (the 2048x1024x512
        // channel's 40 MB y slabs allow two: 33 900 against 38 200 MLUPS), so
        // the slabs are then cut into row windows (there: four of 256 rows,
        // 20 to a chunk, 42 100-42 200 MLUPS in FP32).  On
        // the 192x768x384 channel this picks 11.2 MB slabs, 18 to a chunk, in
        // both FP64 (y: 20 650-20 700 MLUPS against 20 000-20 090 for z
        // slabs, 8 to a chunk) and FP32 (z: about 42 000).
        // LBM_BLOCK_AXIS=z|y sets the axis, LBM_BLOCK_PLANES=K the chunk (0:
        // off), LBM_BLOCK_ROWS=W row windows of W rows (0: whole slabs).  Not
        // by default for 2-byte storage, whose step is bound by arithmetic
        // (FP16: measured 66 300 MLUPS blocked, 32 to a chunk, against 72 000
        // single-step).
        const BlockChoice bc =
            block_choice(h->g, h->elem, h->elem != 2, BLOCK_RING_BYTES, BLOCK_MIN_SLABS, getenv("LBM_BLOCK_AXIS"),
                         getenv("LBM_BLOCK_PLANES"), getenv("LBM_BLOCK_ROWS"));
        if (bc.bad) {
            delete h;
            return fail(LBM_ERR_ARG, "lbm_create: invalid value of %s", bc.bad);
        }
        h->baxis = bc.axis;
        h->bK = bc.K;
        h->bW = bc.W;
        // LBM_BLOCK_FORCED=1: pairs under a force evaluated per node too.
        // Off by default: on the 192x768x384 channel such pairs took 3.33 ms
        // per FP64 step against 3.10 ms for single steps (the per-node force
        // costs the latency-bound sweep what the HBM-bound step hides).
        const char *kf = getenv("LBM_BLOCK_FORCED");
        h->bforced = (kf && strcmp(kf, "1") == 0) ? 1 : 0;
        // LBM_DEFER=0: collision_MRT launches at once (no per-call pairing)
        const char *kd = getenv("LBM_DEFER");
        if (kd && strcmp(kd, "0") != 0 && strcmp(kd, "1") != 0) {
            delete h;
            return fail(LBM_ERR_ARG, "lbm_create: invalid value of LBM_DEFER (0 or 1)");
        }
        h->defer = (kd && strcmp(kd, "0") == 0) ? 0 : 1;
    }
    fields_from_cur(h);
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
    const size_t part = 2 * RED_BLOCKS + (size_t)PLANE_CHUNKS * NPLANE * c.lx;
    if (hipMalloc((void **)&h->partials, part * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&h->red_out, (2 + NPLANE * (size_t)c.lx) * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&h->stat_acc, NPLANE * (size_t)c.lx * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&h->errbits, sizeof(unsigned long long)) != hipSuccess)
        return cleanup(fail(LBM_ERR_NOMEM, "hipMalloc of reduction scratch failed"));
    e = hipMemsetAsync(h->stat_acc, 0, NPLANE * (size_t)c.lx * sizeof(double), h->stream);
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
    if (h->pstream) // particle work an error return left unjoined
        (void)hipStreamSynchronize(h->pstream);
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
    if (h->fstat_acc)
        (void)hipFree(h->fstat_acc);
    if (h->errbits)
        (void)hipFree(h->errbits);
    if (h->stage)
        (void)hipFree(h->stage);
    if (h->part) {
        for (void *p : h->part->allocs)
            (void)hipFree(p);
        delete h->part;
    }
    if (h->mlp)
        (void)hipFree(h->mlp);
    if (h->ftab)
        (void)hipFree(h->ftab);
    if (h->bring)
        (void)hipFree(h->bring);
    if (h->bhead)
        (void)hipFree(h->bhead);
    for (hipEvent_t e : h->pev)
        if (e)
            (void)hipEventDestroy(e);
    if (h->pstream)
        (void)hipStreamDestroy(h->pstream);
    if (h->ev[0])
        (void)hipEventDestroy(h->ev[0]);
    if (h->ev[1])
        (void)hipEventDestroy(h->ev[1]);
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
    return LBM_OK;
}

// Launches the collisions lbm_collision_mrt deferred, as single steps, and
// leaves the state the calls one by one would have left: the first step, the
// fields of a macrovar that followed it, the second step (whose macrovar has
// not come, or the pair would have run).
static int flush_pending(lbm_channel *h)
{
    const int n = h->pend;
    if (n == 0)
        return LBM_OK;
    const bool mac = h->pend_mac;
    h->pend = 0;
    h->pend_mac = false;
    int rc = launch_step(h);
    if (rc)
        return rc;
    if (n == 1 && !mac)
        return LBM_OK; // fields FROM_PREV, as launch_step left them
    fields_from_cur(h);
    h->rho_off = 0.0;
    return n == 2 ? launch_step(h) : LBM_OK;
}

// Entry with no deferred work flushed: lbm_collision_mrt and lbm_macrovar.
#define CHECK_H_KEEP(h)                                                  \
    do {                                                                 \
        if (!(h))                                                        \
            return fail(LBM_ERR_ARG, "%s: null handle", __func__);       \
        g_err[0] = 0;                                                    \
        HIP_TRY(hipSetDevice((h)->cfg.device));                          \
    } while (0)

// Every other entry point first launches deferred collisions, so it sees the
// state the calls one by one leave.
#define CHECK_H(h)                                                       \
    do {                                                                 \
        CHECK_H_KEEP(h);                                                 \
        const int frc_ = flush_pending(h);                               \
        if (frc_)                                                        \
            return frc_;                                                 \
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
    // the fields keep the force they were formed with (fm) until the next
    // macrovar; a fresh handle has formed none yet
    h->fc.fx = fx;
    h->fc.fy = fy;
    h->fc.fz = fz;
    h->fc.tab = nullptr;
    h->fc.a = h->fc.b = h->fc.c = 0.0;
    if (h->fresh)
        h->fm = h->fc;
    return LBM_OK;
}

int lbm_forcingp(lbm_channel *h, int istep, double force_in_y)
{
    CHECK_H(h);
    if (h->part)
        return fail(LBM_ERR_STATE, "lbm_forcingp: not for particle-laden channels (Channel-Flow/main.f90:148)");
    const uint32_t lx = h->g.lx, ly = h->g.ly, lz = h->g.lz;
    const double pi2 = 2.0 * (4.0 * atan(1.0)); // var_inc.f90:71-72
    const double beta9 = 3.0, gamma9 = 2.0, phase9 = 0.25, Tpd = 2000.;
    const int ixs0 = 2;
    const int ihh = (int)(lx / 2) / 2; // lxh/2, var_inc.f90:53
    if (!h->ftab) {
        // factors of one coordinate each (collision.f90:561-599, one rank);
        // computed with the C library's sin and cos, as the oracle does
        std::vector<double> t(3 * (size_t)lx + 4 * (size_t)ly + 2 * (size_t)lz, 0.0);
        for (int band = 1; band <= 2; band++) {
            const int ixs = band == 1 ? ixs0 : (int)lx - ixs0 - ihh; // the second band is written last
            for (int i = 1; i <= ihh; i++) {
                const double x9 = pi2 * ((double)i - 0.5) / (double)ihh;
                double *xt = &t[3 * (size_t)(i + ixs - 1)];
                xt[0] = (double)band;
                xt[1] = 1. - cos(x9);
                xt[2] = sin(x9);
            }
        }
        for (uint32_t j = 1; j <= ly; j++) {
            const double y1 = pi2 * ((double)j - 0.5) / (double)ly;
            const double y2 = pi2 * (((double)j - 0.5) / (double)ly + phase9);
            double *yt = &t[3 * (size_t)lx + 4 * (size_t)(j - 1)];
            yt[0] = cos(beta9 * y1);
            yt[1] = sin(beta9 * y1);
            yt[2] = cos(beta9 * y2);
            yt[3] = sin(beta9 * y2);
        }
        for (uint32_t k = 1; k <= lz; k++) {
            const double z9 = pi2 * ((double)k - 0.5) / (double)lz;
            double *zt = &t[3 * (size_t)lx + 4 * (size_t)ly + 2 * (size_t)(k - 1)];
            zt[0] = cos(gamma9 * z9);
            zt[1] = sin(gamma9 * z9);
        }
        hipError_t e = hipMalloc((void **)&h->ftab, t.size() * sizeof(double));
        if (e != hipSuccess)
            return fail(LBM_ERR_NOMEM, "hipMalloc of the FORCINGP tables failed: %s", hipGetErrorString(e));
        HIP_TRY(hipMemcpyAsync(h->ftab, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    // collision.f90:541 and the amplitude factors of :571-576
    const double Amp0 = 40.00 * beta9 / (double)ly * sin(pi2 * (double)istep / Tpd);
    h->fc.fx = 0.0;
    h->fc.fy = force_in_y;
    h->fc.fz = 0.0;
    h->fc.tab = h->ftab;
    h->fc.a = force_in_y * 0.5 * Amp0 * (double)ihh;
    h->fc.b = Amp0 * (double)ly / beta9;
    h->fc.c = force_in_y * 0.5 * Amp0 * (double)lz / gamma9;
    if (h->fresh)
        h->fm = h->fc;
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

namespace {

constexpr int SEP_MAX_TERMS = 64;
struct SepTerms {
    int n;
    int comp[SEP_MAX_TERMS];
};

// rho = 0 and u_c = sum over the terms k of component c of
// planes[k][row] * profiles[k][ix] (row = iy + ly*iz), in increasing k from 0.
__global__ void k_separable(Geom g, SepTerms t, const double *__restrict__ planes,
                            const double *__restrict__ profiles, double *rho, double *ux, double *uy, double *uz)
{
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.N)
        return;
    const uint64_t row = n / g.lx, ix = n - row * g.lx;
    const uint64_t rows = g.N / g.lx;
    double u[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < t.n; k++)
        u[t.comp[k]] = u[t.comp[k]] + planes[(uint64_t)k * rows + row] * profiles[(uint64_t)k * g.lx + ix];
    rho[n] = 0.0;
    ux[n] = u[0];
    uy[n] = u[1];
    uz[n] = u[2];
}

} // namespace

int lbm_set_velocity_separable(lbm_channel *h, int nterms, const int *comp, const double *planes,
                               const double *profiles)
{
    CHECK_H(h);
    if (nterms < 1 || nterms > SEP_MAX_TERMS || !comp || !planes || !profiles)
        return fail(LBM_ERR_ARG, "lbm_set_velocity_separable: need 1 <= nterms <= %d and non-null arrays",
                    SEP_MAX_TERMS);
    SepTerms t;
    t.n = nterms;
    for (int k = 0; k < nterms; k++) {
        if (comp[k] < 0 || comp[k] > 2)
            return fail(LBM_ERR_ARG, "lbm_set_velocity_separable: term %d has component %d (0, 1 or 2)", k, comp[k]);
        t.comp[k] = comp[k];
    }
    int rc = materialize(h);
    if (rc)
        return rc;
    const size_t rows = (size_t)h->g.ly * h->g.lz;
    const size_t pb = (size_t)nterms * rows * sizeof(double), qb = (size_t)nterms * h->g.lx * sizeof(double);
    double *dp = nullptr, *dq = nullptr;
    hipError_t e = hipMalloc((void **)&dp, pb);
    if (e == hipSuccess)
        e = hipMalloc((void **)&dq, qb);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dp, planes, pb, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dq, profiles, qb, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_separable, dim3((unsigned)((h->g.N + 255) / 256)), dim3(256), 0, h->stream, h->g, t, dp,
                           dq, h->mac[0], h->mac[1], h->mac[2], h->mac[3]);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(h->stream);
    if (dp)
        (void)hipFree(dp);
    if (dq)
        (void)hipFree(dq);
    if (e != hipSuccess)
        return fail(LBM_ERR_HIP, "lbm_set_velocity_separable: %s", hipGetErrorString(e));
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

static bool block_ready(const lbm_channel *h);
static int run_pair(lbm_channel *h);

int lbm_collision_mrt(lbm_channel *h)
{
    CHECK_H_KEEP(h);
    // Per-call pairing: a collision that lbm_run could pair is deferred; if
    // macrovar and another collision follow, the two run as one blocked pair
    // at the next macrovar (bitwise the single steps).  Any other call
    // launches what is deferred first.
    if (h->defer && h->pend == 0 && block_ready(h)) {
        h->pend = 1;
        h->pend_mac = false;
        return LBM_OK;
    }
    if (h->pend == 1 && h->pend_mac) {
        h->pend = 2;
        return LBM_OK;
    }
    const int rc = flush_pending(h);
    if (rc)
        return rc;
    return launch_step(h);
}

int lbm_macrovar(lbm_channel *h, double *rho, double *ux, double *uy, double *uz)
{
    CHECK_H_KEEP(h);
    if (h->pend && !rho && !ux && !uy && !uz) {
        if (h->pend == 1) { // fields of the deferred step (again, if twice)
            h->pend_mac = true;
            return LBM_OK;
        }
        h->pend = 0; // collision, macrovar, collision, macrovar: one pair
        h->pend_mac = false;
        return run_pair(h);
    }
    const int frc = flush_pending(h);
    if (frc)
        return frc;
    fields_from_cur(h);
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
    a.f = h->fm;
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
    double sum = 0.0, cnt = 0.0, mean;
    int rc;
    if (h->ms == FROM_CUR && h->rho_off == 0.0) {
        // rho is virtual: reduce macrovar(F) directly and keep the correction
        // as an offset applied by the next collision
        rc = reduce_rho(h, 1, &sum, &cnt);
        if (rc)
            return rc;
        mean = cnt > 0.0 ? sum / cnt : 0.0;
        h->rho_off = mean;
    } else {
        rc = materialize(h);
        if (rc)
            return rc;
        rc = reduce_rho(h, 0, &sum, &cnt);
        if (rc)
            return rc;
        mean = cnt > 0.0 ? sum / cnt : 0.0;
        hipLaunchKernelGGL(k_sub_scalar, dim3((unsigned)((h->g.N + 255) / 256)), dim3(256), 0, h->stream, h->mac[0],
                           h->g.N, mean);
        HIP_TRY(hipGetLastError());
    }
    h->fresh = false;
    if (rhomean)
        *rhomean = mean;
    return LBM_OK;
}

// Can the next two main-loop steps (collision_MRT + macrovar, twice) run as
// one blocked sweep?  Fused fields of the current lattice with no pending
// density offset, one uniform force, no particles, no closure.
static bool block_ready(const lbm_channel *h)
{
    return h->bK > 0 && h->ms == FROM_CUR && h->rho_off == 0.0 && !h->part && !h->mlp && h->fc.tab == nullptr &&
           same_force(h->fm, h->fc);
}

// The force of one step of a pair: collision force fc, fields formed with fm
// (per node when the collision would run mode 16, as in launch_step).
static StepForce step_force_of(const lbm_channel *h, const ForceSpec &fc, const ForceSpec &fm)
{
    StepForce f;
    f.r = rates_of(h);
    f.r.fx = fc.fx;
    f.r.fy = fc.fy;
    f.r.fz = fc.fz;
    f.r = with_f32(f.r);
    f.fc = fc;
    f.fm = fm;
    f.per_node = (fc.tab != nullptr || !same_force(fm, fc)) ? 1 : 0;
    return f;
}

// Two main-loop steps in one blocked sweep (blocked.h): step t with force f0,
// step t+1 with f1; the handle's force (h->fc) is then step t+1's.
static int run_pair_with(lbm_channel *h, const StepForce &f0, const StepForce &f1)
{
    if (!h->bring) {
        const size_t rb = block_ring_elems(h->g, h->baxis, h->bK, h->bW) * h->elem,
                     hb = block_head_elems(h->g, h->baxis, h->bW) * h->elem;
        hipError_t e = hipMalloc(&h->bring, rb);
        if (e == hipSuccess)
            e = hipMalloc(&h->bhead, hb);
        if (e != hipSuccess) {
            if (h->bring)
                (void)hipFree(h->bring);
            h->bring = nullptr;
            return fail(LBM_ERR_NOMEM, "hipMalloc of the blocked-sweep ring (%zu + %zu bytes) failed: %s", rb, hb,
                        hipGetErrorString(e));
        }
    }
    const auto launches =
        block_schedule(h->g, h->baxis, f0, f1, h->bK, h->bW, h->lat[h->cur], h->lat[1 - h->cur], h->bring,
                       h->bhead);
    for (const BlockArgs &a : launches) {
        int rc = dispatch_err(PREC_CALL(h, step_planes, a, h->baxis, h->stream), "collision (blocked sweep)");
        if (rc)
            return rc;
    }
    h->cur = 1 - h->cur;
    h->have_prev = false; // the other lattice holds the state two steps back
    h->fresh = false;
    fields_from_cur(h); // macrovar
    h->rho_off = 0.0;
    return LBM_OK;
}

// Two main-loop steps under the handle's uniform force.
static int run_pair(lbm_channel *h)
{
    const StepForce f = step_force_of(h, h->fc, h->fm);
    return run_pair_with(h, f, f);
}

// The force the main loop sets before the collision of step istep
// (Channel-Flow/main.f90:148-153 with the perturbation block switched on):
// FORCINGP while istep < npforcing, FORCING at istep == npforcing.
static int main_loop_force(lbm_channel *h, int istep, int npforcing, double force_in_y, double fy_forcing)
{
    if (istep < npforcing)
        return lbm_forcingp(h, istep, force_in_y);
    if (istep == npforcing)
        return lbm_set_force(h, 0.0, fy_forcing, 0.0);
    return LBM_OK;
}

int lbm_run_forcingp(lbm_channel *h, int istep0, int nsteps, int npforcing, double force_in_y, double fy_forcing)
{
    CHECK_H(h);
    if (nsteps < 0)
        return fail(LBM_ERR_ARG, "lbm_run_forcingp: nsteps %d < 0", nsteps);
    int s = 0;
    while (s < nsteps) {
        const int istep = istep0 + s + 1;
        int rc = main_loop_force(h, istep, npforcing, force_in_y, fy_forcing);
        if (rc)
            return rc;
        // a pair as one blocked sweep: fused fields of the current lattice,
        // no density offset, no particles or closure, and one uniform force
        // (or, with bforced, any forces, evaluated per node)
        bool paired = false;
        if (nsteps - s >= 2 && h->bK > 0 && h->ms == FROM_CUR && h->rho_off == 0.0 && !h->part && !h->mlp) {
            const ForceSpec fc0 = h->fc, fm0 = h->fm;
            rc = main_loop_force(h, istep + 1, npforcing, force_in_y, fy_forcing);
            if (rc)
                return rc;
            // step t+1's fields are macrovar of step t, formed with fc0
            const StepForce f0 = step_force_of(h, fc0, fm0), f1 = step_force_of(h, h->fc, fc0);
            if (h->bforced || (!f0.per_node && !f1.per_node)) {
                rc = run_pair_with(h, f0, f1);
                if (rc)
                    return rc;
                s += 2;
                paired = true;
            } else {
                h->fc = fc0; // step t alone; step t+1 sets its force again
                h->fm = fm0;
            }
        }
        if (!paired) {
            rc = launch_step(h);
            if (rc)
                return rc;
            fields_from_cur(h); // macrovar
            h->rho_off = 0.0;
            s += 1;
        }
    }
    return LBM_OK;
}

int lbm_run(lbm_channel *h, int nsteps)
{
    CHECK_H(h);
    if (nsteps < 0)
        return fail(LBM_ERR_ARG, "lbm_run: nsteps %d < 0", nsteps);
    int s = 0;
    while (nsteps - s >= 2 && block_ready(h)) {
        int rc = run_pair(h);
        if (rc)
            return rc;
        s += 2;
    }
    for (; s < nsteps; s++) {
        int rc = launch_step(h);
        if (rc)
            return rc;
        fields_from_cur(h); // macrovar
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
    a.solid = solid_of(h);
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

int lbm_mass_sum_noalloc(lbm_channel *h, double *mass_dev)
{
    CHECK_H(h);
    if (!mass_dev)
        return fail(LBM_ERR_ARG, "lbm_mass_sum_noalloc: null output");
    // a kernel store through a host pointer would fault the GPU: accept only
    // device memory of this handle's GPU
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, mass_dev) != hipSuccess || at.type != hipMemoryTypeDevice ||
        at.device != h->cfg.device) {
        (void)hipGetLastError();
        return fail(LBM_ERR_ARG, "lbm_mass_sum_noalloc: output %p is not device memory of device %d", (void *)mass_dev,
                    h->cfg.device);
    }
    double *part = g_scratch_ptr[h->cfg.device];
    if (!part)
        return fail(LBM_ERR_STATE, "lbm_mass_sum_noalloc: scratch of device %d not set up", h->cfg.device);
    SumArgs a;
    a.g = h->g;
    a.L = h->lat[h->cur];
    a.rho = nullptr;
    a.solid = solid_of(h);
    a.partials = part;
    // the same grid as lbm_mass_sum, so the two sums are equal bit for bit
    const uint32_t grid = (uint32_t)std::min<uint64_t>(RED_BLOCKS, (h->g.N + 255) / 256);
    if (grid > NOALLOC_PARTIALS) // cannot happen while RED_BLOCKS bounds the grid
        return fail(LBM_ERR_ARG, "lbm_mass_sum_noalloc: %u partials exceed the %u-double scratch", grid,
                    NOALLOC_PARTIALS);
    int rc = dispatch_err(PREC_CALL(h, mass_partials, a, grid, h->stream), "mass reduction");
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_sum_columns, dim3(1), dim3(64), 0, h->stream, (const double *)part, grid, 1u, mass_dev,
                       (double *)nullptr);
    HIP_TRY(hipGetLastError());
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
    std::vector<double> s(NPLANE * (size_t)h->g.lx);
    HIP_TRY(hipMemcpyAsync(s.data(), h->red_out + 2, s.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    stat_columns(h->g.lx, (double)h->g.ly * (double)h->g.lz, s.data(), 1.0, ustar, ystar, out);
    return LBM_OK;
}

namespace {

// diag's block partials (q*grid + block) to totals in out[0..DIAG_Q): the sums
// in increasing block order from 0.0, the maxima merged in block order.
__global__ void k_diag_final(const double *__restrict__ part, uint32_t grid, double *out)
{
    if (threadIdx.x != 0)
        return;
    double t[DIAG_Q];
    for (int q = 0; q < 7; q++) {
        double s = 0.0;
        for (uint32_t b = 0; b < grid; b++)
            s += part[(uint64_t)q * grid + b];
        t[q] = s;
    }
    double v = part[7 * (uint64_t)grid], vi = part[8 * (uint64_t)grid];
    double rx = part[9 * (uint64_t)grid], rn = part[10 * (uint64_t)grid];
    for (uint32_t b = 1; b < grid; b++) {
        const double ov = part[7 * (uint64_t)grid + b], oi = part[8 * (uint64_t)grid + b];
        if (ov > v || (ov == v && oi >= 0.0 && (vi < 0.0 || oi < vi))) {
            v = ov;
            vi = oi;
        }
        const double ox = part[9 * (uint64_t)grid + b], on = part[10 * (uint64_t)grid + b];
        rx = ox > rx ? ox : rx;
        rn = on < rn ? on : rn;
    }
    t[7] = v;
    t[8] = vi;
    t[9] = rx;
    t[10] = rn;
    for (int q = 0; q < DIAG_Q; q++)
        out[q] = t[q];
}

// vortcalc (Channel-Flow/saveload.f90:3929-4023) from the field arrays.
__global__ void k_vortcalc(Geom g, const double *__restrict__ ux, const double *__restrict__ uy,
                           const double *__restrict__ uz, const int *__restrict__ solid, const double *__restrict__ omg,
                           double *ox, double *oy, double *oz)
{
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.N)
        return;
    if (solid && solid[n] >= 0) { // twice the particle's angular velocity
        const double *w = omg + 3 * (uint64_t)solid[n];
        ox[n] = 2.0 * w[0];
        oy[n] = 2.0 * w[1];
        oz[n] = 2.0 * w[2];
        return;
    }
    const uint64_t row = n / g.lx;
    const uint32_t i = (uint32_t)(n - row * g.lx);
    const uint32_t k = (uint32_t)(row / g.ly), j = (uint32_t)(row - (uint64_t)k * g.ly);
    const uint64_t plane = (uint64_t)g.lx * g.ly;
    double pwx, pvx;
    if (i == 0) { // wall forms of :3970-3975
        pwx = (3.0 * uz[n] + uz[n + 1]) / 3.0;
        pvx = (3.0 * uy[n] + uy[n + 1]) / 3.0;
    } else if (i == g.lx - 1) {
        pwx = -(3.0 * uz[n] + uz[n - 1]) / 3.0;
        pvx = -(3.0 * uy[n] + uy[n - 1]) / 3.0;
    } else {
        pwx = (uz[n + 1] - uz[n - 1]) / 2.0;
        pvx = (uy[n + 1] - uy[n - 1]) / 2.0;
    }
    // periodic images in y and z (exchng8, :4027-4054, on one rank)
    const uint64_t yp = j + 1 == g.ly ? n - (uint64_t)(g.ly - 1) * g.lx : n + g.lx;
    const uint64_t ym = j == 0 ? n + (uint64_t)(g.ly - 1) * g.lx : n - g.lx;
    const uint64_t zp = k + 1 == g.lz ? n - (uint64_t)(g.lz - 1) * plane : n + plane;
    const uint64_t zm = k == 0 ? n + (uint64_t)(g.lz - 1) * plane : n - plane;
    const double pwy = (uz[yp] - uz[ym]) / 2.0;
    const double puy = (ux[yp] - ux[ym]) / 2.0;
    const double puz = (ux[zp] - ux[zm]) / 2.0;
    const double pvz = (uy[zp] - uy[zm]) / 2.0;
    ox[n] = pwy - pvz;
    oy[n] = puz - pwx;
    oz[n] = pvx - puy;
}

} // namespace

int lbm_diag(lbm_channel *h, double ustar, double *out)
{
    CHECK_H(h);
    if (!out)
        return fail(LBM_ERR_ARG, "lbm_diag: null output");
    if (h->ms == FROM_PREV) {
        const int rc = materialize(h);
        if (rc)
            return rc;
    }
    DiagArgs a;
    a.g = h->g;
    a.L = h->lat[h->cur];
    a.rho = h->mac[0];
    a.ux = h->mac[1];
    a.uy = h->mac[2];
    a.uz = h->mac[3];
    a.solid = solid_of(h);
    a.f = h->fm;
    a.rho_off = h->rho_off;
    a.partials = h->partials;
    // DIAG_Q partials per block; the scratch holds at least 16 384 doubles
    const uint32_t grid = (uint32_t)std::min<uint64_t>(1024, (h->g.N + 255) / 256);
    int rc = dispatch_err(PREC_CALL(h, diag_partials, a, h->ms == ARRAYS ? 0 : 1, grid, h->stream), "diag");
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, h->stream, (const double *)h->partials, grid,
                       h->red_out + 2);
    HIP_TRY(hipGetLastError());
    double t[DIAG_Q];
    HIP_TRY(hipMemcpyAsync(t, h->red_out + 2, sizeof(t), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    // saveload.f90:1627-1640, in its order
    const double rn = t[6];
    double um = t[0] / rn, vm = t[1] / rn, wm = t[2] / rn;
    const double ur = sqrt(t[3] / rn - um * um), vr = sqrt(t[4] / rn - vm * vm), wr = sqrt(t[5] / rn - wm * wm);
    const long long idx = t[8] < 0.0 ? -1 : (long long)t[8];
    out[0] = t[7];
    out[1] = idx < 0 ? 0.0 : (double)(idx % h->g.lx + 1);
    out[2] = idx < 0 ? 0.0 : (double)((idx / h->g.lx) % h->g.ly + 1);
    out[3] = idx < 0 ? 0.0 : (double)(idx / ((long long)h->g.lx * h->g.ly) + 1);
    out[4] = um / ustar;
    out[5] = vm / ustar;
    out[6] = wm / ustar;
    out[7] = ur / ustar;
    out[8] = vr / ustar;
    out[9] = wr / ustar;
    out[10] = 1. - rn / ((double)h->g.lx * (double)h->g.ly * (double)h->g.lz);
    out[11] = t[9];
    out[12] = t[10];
    return LBM_OK;
}

int lbm_vortcalc(lbm_channel *h, double *ox, double *oy, double *oz)
{
    CHECK_H(h);
    if (!ox || !oy || !oz)
        return fail(LBM_ERR_ARG, "lbm_vortcalc: null output");
    int rc = materialize(h);
    if (rc)
        return rc;
    const size_t bytes = h->g.N * sizeof(double);
    double *d = nullptr;
    hipError_t e = hipMalloc((void **)&d, 3 * bytes);
    if (e != hipSuccess)
        return fail(LBM_ERR_NOMEM, "lbm_vortcalc: hipMalloc of %zu bytes failed: %s", 3 * bytes, hipGetErrorString(e));
    const double *omg = h->part ? h->part->ps.W : nullptr;
    hipLaunchKernelGGL(k_vortcalc, dim3((unsigned)((h->g.N + 255) / 256)), dim3(256), 0, h->stream, h->g,
                       (const double *)h->mac[1], (const double *)h->mac[2], (const double *)h->mac[3], solid_of(h),
                       omg, d, d + h->g.N, d + 2 * h->g.N);
    e = hipGetLastError();
    double *dst[3] = {ox, oy, oz};
    for (int k = 0; k < 3 && e == hipSuccess; k++)
        e = hipMemcpyAsync(dst[k], d + (size_t)k * h->g.N, bytes, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess)
        return fail(LBM_ERR_HIP, "lbm_vortcalc: %s", hipGetErrorString(e));
    return LBM_OK;
}

int lbm_dissipation(lbm_channel *h, double tau, double ystar, double *out)
{
    CHECK_H(h);
    if (!out || !(tau > 0.0))
        return fail(LBM_ERR_ARG, "lbm_dissipation: need an output and tau > 0");
    if (h->ms == FROM_PREV) { // f is the current lattice's, the fields the previous macrovar's
        const int rc = materialize(h);
        if (rc)
            return rc;
    }
    const lbm_config &c = h->cfg;
    DissArgs a;
    a.g = h->g;
    a.L = h->lat[h->cur];
    a.rho = h->mac[0];
    a.ux = h->mac[1];
    a.uy = h->mac[2];
    a.uz = h->mac[3];
    a.solid = solid_of(h);
    a.f = h->fm;
    a.rho_off = h->rho_off;
    a.s1 = c.s1;
    a.s9 = c.s9;
    a.coef2 = c.coef2;
    a.coef3 = c.coef3;
    a.coef5 = c.coef5;
    a.ww1 = c.ww1;
    a.ww2 = c.ww2;
    a.c15tau = 1.5 / tau;
    a.partials = h->partials + 2 * RED_BLOCKS;
    const uint32_t rows = h->g.ly * h->g.lz;
    const uint32_t chunks = std::min<uint32_t>(PLANE_CHUNKS, rows);
    int rc = dispatch_err(PREC_CALL(h, dissipation_partials, a, h->ms == ARRAYS ? 0 : 1, chunks, h->stream),
                          "dissipation");
    if (rc)
        return rc;
    const uint32_t width = NDISS * h->g.lx;
    hipLaunchKernelGGL(k_sum_columns, dim3((width + 255) / 256), dim3(256), 0, h->stream, a.partials, chunks, width,
                       h->red_out + 2, (double *)nullptr);
    HIP_TRY(hipGetLastError());
    std::vector<double> t(width);
    HIP_TRY(hipMemcpyAsync(t.data(), h->red_out + 2, width * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const int lx = h->g.lx;
    for (int i = 1; i <= lx; i++) { // saveload.f90:4857-4875
        const double nf = t[4 * (size_t)lx + i - 1];
        double yplus = ((double)i - 0.5) / ystar;
        if (i > lx / 2)
            yplus = ((double)(lx - i) + 0.5) / ystar;
        double *o = out + 6 * (size_t)(i - 1);
        o[0] = (double)i - 0.5;
        o[1] = yplus;
        for (int q = 0; q < 4; q++)
            o[2 + q] = t[(size_t)q * lx + i - 1] / nf;
    }
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
    std::vector<double> s(NPLANE * (size_t)h->g.lx);
    HIP_TRY(hipMemcpyAsync(s.data(), h->stat_acc, s.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    stat_columns(h->g.lx, (double)h->g.ly * (double)h->g.lz, s.data(), (double)h->stat_n, ustar, ystar, out);
    return LBM_OK;
}

int lbm_stats_accumulate_filtered(lbm_channel *h)
{
    CHECK_H(h);
    if (h->g.lx % 2 || h->g.ly % 2 || h->g.lz % 2)
        return fail(LBM_ERR_ARG, "lbm_stats_accumulate_filtered: the 2x2x2 filter needs even lx, ly, lz (%u, %u, %u)",
                    h->g.lx, h->g.ly, h->g.lz);
    if (h->part)
        return fail(LBM_ERR_STATE, "lbm_stats_accumulate_filtered: not defined with particles (filter cells would mix "
                                   "fluid and solid nodes)");
    if (!h->fstat_acc) {
        const uint64_t n = NPLANE * (uint64_t)(h->g.lx / 2);
        if (hipMalloc((void **)&h->fstat_acc, n * sizeof(double)) != hipSuccess)
            return fail(LBM_ERR_NOMEM, "hipMalloc of the filtered statistics accumulator failed");
        hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->fstat_acc, n, 0.0);
        HIP_TRY(hipGetLastError());
        h->fstat_n = 0;
    }
    int rc = filtered_plane_sums(h);
    if (rc)
        return rc;
    h->fstat_n++;
    return LBM_OK;
}

int lbm_stats_read_filtered(lbm_channel *h, double ustar, double ystar, double *out, long long *nsamples)
{
    CHECK_H(h);
    if (!out)
        return fail(LBM_ERR_ARG, "lbm_stats_read_filtered: null output");
    if (nsamples)
        *nsamples = h->fstat_n;
    if (h->fstat_n == 0)
        return fail(LBM_ERR_STATE, "lbm_stats_read_filtered: no samples accumulated");
    const uint32_t lxc = h->g.lx / 2;
    std::vector<double> s(NPLANE * (size_t)lxc);
    HIP_TRY(hipMemcpyAsync(s.data(), h->fstat_acc, s.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    stat_columns(lxc, (double)(h->g.ly / 2) * (double)(h->g.lz / 2), s.data(), (double)h->fstat_n, ustar, ystar,
                 out);
    return LBM_OK;
}

int lbm_stats_reset(lbm_channel *h)
{
    CHECK_H(h);
    const uint64_t n = NPLANE * (uint64_t)h->g.lx;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->stat_acc, n, 0.0);
    HIP_TRY(hipGetLastError());
    h->stat_n = 0;
    if (h->fstat_acc) {
        const uint64_t nf = NPLANE * (uint64_t)(h->g.lx / 2);
        hipLaunchKernelGGL(k_fill, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, h->stream, h->fstat_acc, nf,
                           0.0);
        HIP_TRY(hipGetLastError());
    }
    h->fstat_n = 0;
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
    // one timed unit per single step, or per blocked pair of steps; the
    // events are destroyed on every return path
    struct Events {
        std::vector<hipEvent_t> e;
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x)
                    (void)hipEventDestroy(x);
        }
    } evs;
    evs.e.assign(2 * (size_t)nsteps, nullptr);
    std::vector<hipEvent_t> &ev = evs.e;
    for (auto &e : ev)
        HIP_TRY(hipEventCreate(&e));
    int rc = LBM_OK;
    int units = 0;
    for (int s = 0; s < nsteps && rc == LBM_OK; units++) {
        HIP_TRY(hipEventRecord(ev[2 * units], h->stream));
        if (nsteps - s >= 2 && block_ready(h)) {
            rc = run_pair(h);
            s += 2;
        } else {
            rc = launch_step(h);
            fields_from_cur(h);
            h->rho_off = 0.0;
            s += 1;
        }
        HIP_TRY(hipEventRecord(ev[2 * units + 1], h->stream));
    }
    if (rc == LBM_OK) {
        HIP_TRY(hipEventSynchronize(ev[2 * units - 1]));
        double sum = 0.0;
        for (int u = 0; u < units; u++) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, ev[2 * u], ev[2 * u + 1]));
            sum += ms;
        }
        float tot = 0.f;
        HIP_TRY(hipEventElapsedTime(&tot, ev[0], ev[2 * units - 1]));
        if (avg_ms)
            *avg_ms = sum / nsteps; // device time per step
        if (total_ms)
            *total_ms = tot;
    }
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
    if (h->bring)
        b += (block_ring_elems(h->g, h->baxis, h->bK, h->bW) + block_head_elems(h->g, h->baxis, h->bW)) * h->elem;
    b += (2 * RED_BLOCKS + (size_t)PLANE_CHUNKS * NPLANE * h->g.lx + 2 + 2 * NPLANE * (size_t)h->g.lx) * sizeof(double) + 8;
    *bytes = b;
    return LBM_OK;
}

int lbm_block_info(lbm_channel *h, int *axis, int *slabs)
{
    CHECK_H(h);
    if (axis)
        *axis = h->baxis == 2 ? 'z' : 'y';
    if (slabs)
        *slabs = (int)h->bK;
    return LBM_OK;
}

int lbm_block_rows(lbm_channel *h, int *rows)
{
    CHECK_H(h);
    if (!rows)
        return fail(LBM_ERR_ARG, "lbm_block_rows: null output");
    *rows = (int)h->bW;
    return LBM_OK;
}

// Nodes [n0, n0 + count) of the stored post-collision lattice as FP64,
// population-major (out[i*count + k] = P_i(n0 + k)).
static int download_post_range(lbm_channel *h, uint64_t n0, uint64_t count, double *out)
{
    const uint64_t N = h->g.N;
    const unsigned char *L = (const unsigned char *)h->lat[h->cur];
    if (h->prec == LBM_FP64) {
        for (int i = 0; i < 19; i++)
            HIP_TRY(hipMemcpyAsync(out + (size_t)i * count, L + ((size_t)i * N + n0) * sizeof(double),
                                   count * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        return LBM_OK;
    }
    // FP32/FP16/FP16C: widen on the host, population by population
    std::vector<unsigned char> buf(count * h->elem);
    for (int i = 0; i < 19; i++) {
        HIP_TRY(hipMemcpyAsync(buf.data(), L + ((size_t)i * N + n0) * h->elem, count * h->elem,
                               hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        double *o = out + (size_t)i * count;
        if (h->prec == LBM_FP32) {
            const float *v = (const float *)buf.data();
            for (uint64_t n = 0; n < count; n++)
                o[n] = v[n];
        } else if (h->prec == LBM_FP16) {
            const _Float16 *v = (const _Float16 *)buf.data();
            const double sc = i == 0 ? 0.25 : (i <= 6 ? 1.0 / 16.0 : 1.0 / 32.0);
            for (uint64_t n = 0; n < count; n++)
                o[n] = (double)(float)v[n] * sc;
        } else {
            const uint16_t *v = (const uint16_t *)buf.data();
            const float sc = FP16C_UP / fp16_scale_rt(i);
            for (uint64_t n = 0; n < count; n++)
                o[n] = (double)(fp16c_unpack(v[n]) * sc);
        }
    }
    return LBM_OK;
}

int lbm_download_post(lbm_channel *h, double *post)
{
    CHECK_H(h);
    if (!post)
        return fail(LBM_ERR_ARG, "lbm_download_post: null output");
    return download_post_range(h, 0, h->g.N, post);
}

int lbm_download_post_range(lbm_channel *h, long long first, long long count, double *post)
{
    CHECK_H(h);
    if (!post)
        return fail(LBM_ERR_ARG, "lbm_download_post_range: null output");
    if (first < 0 || count < 0 || (unsigned long long)first + (unsigned long long)count > h->g.N)
        return fail(LBM_ERR_ARG, "lbm_download_post_range: nodes [%lld, %lld + %lld) outside the %llu-node lattice",
                    first, first, count, (unsigned long long)h->g.N);
    if (count == 0)
        return LBM_OK;
    return download_post_range(h, (uint64_t)first, (uint64_t)count, post);
}

static int part_alloc(Part *pt, void **p, size_t bytes)
{
    if (hipMalloc(p, bytes) != hipSuccess)
        return fail(LBM_ERR_NOMEM, "hipMalloc of particle state (%zu bytes) failed", bytes);
    pt->allocs.push_back(*p);
    return LBM_OK;
}

int lbm_particles_create(lbm_channel *h, const lbm_particle_config *pc, const double *pos, const double *vel,
                         const double *omg)
{
    CHECK_H(h);
    if (!pc || !pos)
        return fail(LBM_ERR_ARG, "lbm_particles_create: null argument");
    if (h->part)
        return fail(LBM_ERR_STATE, "lbm_particles_create: particles already exist on this handle");
    if (h->fc.tab)
        return fail(LBM_ERR_STATE, "lbm_particles_create: particles need a uniform force (FORCING, not FORCINGP)");
    if (pc->npart < 1 || !(pc->rad > 0.0) || !(pc->rhopart > 0.0) || pc->nave < 0)
        return fail(LBM_ERR_ARG, "lbm_particles_create: need npart >= 1, rad > 0, rhopart > 0, nave >= 0");
    if (!(pc->mingap > 0.0) || !(pc->mingap_w > 0.0) || !(pc->visc >= 0.0) ||
        (pc->visc > 0.0 && !(pc->lub_cut > pc->lub_min && pc->lub_min > 0.0)))
        return fail(LBM_ERR_ARG, "lbm_particles_create: need mingap, mingap_w > 0, visc >= 0 and, with visc > 0, "
                                 "lub_cut > lub_min > 0");
    const double lx = h->g.lx, ly = h->g.ly, lz = h->g.lz;
    for (int p = 0; p < pc->npart; p++) {
        const double *X = pos + 3 * p;
        if (!(X[0] - pc->rad > 0.0 && X[0] + pc->rad < lx) || !(X[1] >= 0.0 && X[1] < ly) ||
            !(X[2] >= 0.0 && X[2] < lz))
            return fail(LBM_ERR_ARG,
                        "lbm_particles_create: particle %d at (%g, %g, %g) must lie inside the channel "
                        "(R < x < lx - R, 0 <= y < ly, 0 <= z < lz)",
                        p, X[0], X[1], X[2]);
    }
    Part *pt = new Part();
    PartGeom &pg = pt->pg;
    PartState &ps = pt->ps;
    memset(&pg, 0, sizeof(pg));
    memset(&ps, 0, sizeof(ps));
    const double pi = 3.14159265358979323846;
    pg.g = h->g;
    pg.np = pc->npart;
    pg.R = pc->rad;
    pg.R2 = pc->rad * pc->rad;
    // cut links: about 5 per surface node; capacity with margin
    const double area = 4.0 * pi * (pc->rad + 1.0) * (pc->rad + 1.0);
    pg.maxlink = pc->maxlink > 0 ? pc->maxlink : (int)(12.0 * area) + 256;
    pg.maxfill = (int)(4.0 * area) + 64;
    pg.volume = 4.0 / 3.0 * pi * pc->rad * pc->rad * pc->rad;
    pg.mass = pc->rhopart * pg.volume;
    pg.inertia = 0.4 * pg.mass * pc->rad * pc->rad;
    pg.rhopart = pc->rhopart;
    pg.mingap = pc->mingap;
    pg.mingap_w = pc->mingap_w;
    pg.stf0 = pc->stf0;
    pg.stf1 = pc->stf1;
    pg.stf0_w = pc->stf0_w;
    pg.stf1_w = pc->stf1_w;
    pg.visc = pc->visc;
    pg.lub_cut = pc->lub_cut;
    pg.lub_min = pc->lub_min;
    pg.pressure_force = pc->pressure_force;
    pg.nave = pc->nave;
    pg.fx = h->fc.fx;
    pg.fy = h->fc.fy;
    pg.fz = h->fc.fz;
    const size_t np = pc->npart, nl = np * pg.maxlink, nf = np * pg.maxfill;
    int rc = LBM_OK;
    auto A = [&](void **p, size_t bytes) {
        if (rc == LBM_OK)
            rc = part_alloc(pt, p, bytes);
    };
    A((void **)&ps.X, 6 * np * sizeof(double)); // X, then the previous positions
    A((void **)&ps.V, 3 * np * sizeof(double));
    A((void **)&ps.W, 3 * np * sizeof(double));
    A((void **)&ps.TH, 3 * np * sizeof(double));
    A((void **)&ps.Aprev, 3 * np * sizeof(double));
    A((void **)&ps.ALprev, 3 * np * sizeof(double));
    A((void **)&ps.FH, 3 * np * sizeof(double));
    A((void **)&ps.TQ, 3 * np * sizeof(double));
    A((void **)&ps.FC, 3 * np * sizeof(double));
    A((void **)&ps.FL, 3 * np * sizeof(double));
    A((void **)&ps.FHprev, 3 * np * sizeof(double));
    A((void **)&ps.TQprev, 3 * np * sizeof(double));
    A((void **)&ps.VWnew, 6 * np * sizeof(double));
    A((void **)&ps.nsteps_done, sizeof(int));
    A((void **)&ps.solid, h->g.N * sizeof(int));
    A((void **)&ps.link_node, nl * sizeof(int));
    A((void **)&ps.link_dir, nl);
    A((void **)&ps.link_q, nl * sizeof(double));
    A((void **)&ps.link_contrib, nl * 6 * sizeof(double));
    A((void **)&ps.link_count, np * sizeof(int));
    A((void **)&ps.fill_node, nf * sizeof(int));
    A((void **)&ps.fill_count, np * sizeof(int));
    A((void **)&ps.overflow, sizeof(int));
    if (rc) {
        for (void *p : pt->allocs)
            (void)hipFree(p);
        delete pt;
        return rc;
    }
    std::vector<double> zero(3 * np, 0.0);
    HIP_TRY(hipMemcpyAsync(ps.X, pos, 3 * np * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(ps.X + 3 * np, pos, 3 * np * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(ps.V, vel ? vel : zero.data(), 3 * np * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(ps.W, omg ? omg : zero.data(), 3 * np * sizeof(double), hipMemcpyHostToDevice, h->stream));
    for (double *d : {ps.TH, ps.Aprev, ps.ALprev, ps.FH, ps.TQ, ps.FC, ps.FL, ps.FHprev, ps.TQprev})
        HIP_TRY(hipMemcpyAsync(d, zero.data(), 3 * np * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemsetAsync(ps.nsteps_done, 0, sizeof(int), h->stream));
    HIP_TRY(hipMemsetAsync(ps.overflow, 0, sizeof(int), h->stream));
    HIP_TRY(hipMemsetAsync(ps.solid, 0xff, h->g.N * sizeof(int), h->stream)); // -1: fluid
    HIP_TRY(hipMemsetAsync(ps.fill_count, 0, np * sizeof(int), h->stream));
    HIP_TRY(hipMemsetAsync(ps.link_count, 0, np * sizeof(int), h->stream));
    h->part = pt;
    h->ave_pending = false;
    rc = dispatch_err(part_classify(pg, ps, 1, h->stream), "beads_links");
    if (rc)
        return rc;
    int ovf = 0;
    HIP_TRY(hipMemcpyAsync(&ovf, ps.overflow, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (ovf)
        return fail(LBM_ERR_ARG, "lbm_particles_create: cut-link list overflow (maxlink %d)", pg.maxlink);
    return LBM_OK;
}

#define CHECK_P(h)                                                                       \
    do {                                                                                 \
        CHECK_H(h);                                                                      \
        if (!(h)->part)                                                                  \
            return fail(LBM_ERR_STATE, "%s: no particles on this handle", __func__);     \
    } while (0)

static int check_overflow(lbm_channel *h, const char *what)
{
    int ovf = 0;
    HIP_TRY(hipMemcpyAsync(&ovf, h->part->ps.overflow, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (ovf & 2)
        return fail(LBM_ERR_STATE, "%s: a particle's position or velocity became non-finite (the run diverged)",
                    what);
    if (ovf)
        return fail(LBM_ERR_STATE, "%s: a per-particle link or refill list overflowed (maxlink %d, maxfill %d)",
                    what, h->part->pg.maxlink, h->part->pg.maxfill);
    return LBM_OK;
}

int lbm_beads_links(lbm_channel *h)
{
    CHECK_P(h);
    return dispatch_err(part_classify(h->part->pg, h->part->ps, 0, h->stream), "beads_links");
}

int lbm_beads_filling(lbm_channel *h)
{
    CHECK_P(h);
    Part *pt = h->part;
    hipError_t e = h->prec == LBM_FP64   ? PartLaunch<0>::refill(pt->pg, pt->ps, h->lat[h->cur], h->stream)
                   : h->prec == LBM_FP32 ? PartLaunch<1>::refill(pt->pg, pt->ps, h->lat[h->cur], h->stream)
                   : h->prec == LBM_FP16 ? PartLaunch<2>::refill(pt->pg, pt->ps, h->lat[h->cur], h->stream)
                                         : PartLaunch<3>::refill(pt->pg, pt->ps, h->lat[h->cur], h->stream);
    return dispatch_err(e, "beads_filling");
}

int lbm_beads_collision(lbm_channel *h)
{
    CHECK_P(h);
    Part *pt = h->part;
    hipError_t e = h->prec == LBM_FP64   ? PartLaunch<0>::ibb(pt->pg, pt->ps, h->lat[h->cur], h->stream)
                   : h->prec == LBM_FP32 ? PartLaunch<1>::ibb(pt->pg, pt->ps, h->lat[h->cur], h->stream)
                   : h->prec == LBM_FP16 ? PartLaunch<2>::ibb(pt->pg, pt->ps, h->lat[h->cur], h->stream)
                                         : PartLaunch<3>::ibb(pt->pg, pt->ps, h->lat[h->cur], h->stream);
    int rc = dispatch_err(e, "beads_collision");
    if (rc)
        return rc;
    return dispatch_err(part_sum(pt->pg, pt->ps, h->stream), "beads_collision sum");
}

int lbm_beads_move(lbm_channel *h)
{
    CHECK_P(h);
    Part *pt = h->part;
    pt->pg.fx = h->fc.fx;
    pt->pg.fy = h->fc.fy;
    pt->pg.fz = h->fc.fz;
    return dispatch_err(part_move(pt->pg, pt->ps, h->stream), "beads_move");
}

int lbm_run_particles(lbm_channel *h, int istep0, int nsteps)
{
    CHECK_P(h);
    if (nsteps < 0)
        return fail(LBM_ERR_ARG, "lbm_run_particles: nsteps %d < 0", nsteps);
    // The per-particle force sums and beads_move need only the bounce-back's
    // link contributions, and the collision touches no particle state, so
    // they run on a second stream beside the collision; beads_links, which
    // rewrites the node flags the collision reads, waits for both.  Streams
    // of the default priority.
    if (!h->pstream) {
        HIP_TRY(hipStreamCreateWithFlags(&h->pstream, hipStreamNonBlocking));
        for (hipEvent_t &e : h->pev)
            HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    Part *pt = h->part;
    const int nave = pt->pg.nave;
    for (int s = 0; s < nsteps; s++) {
        // beads_collision: the bounce-back here, its sums below
        hipError_t e = h->prec == LBM_FP64   ? PartLaunch<0>::ibb(pt->pg, pt->ps, h->lat[h->cur], h->stream)
                       : h->prec == LBM_FP32 ? PartLaunch<1>::ibb(pt->pg, pt->ps, h->lat[h->cur], h->stream)
                       : h->prec == LBM_FP16 ? PartLaunch<2>::ibb(pt->pg, pt->ps, h->lat[h->cur], h->stream)
                                             : PartLaunch<3>::ibb(pt->pg, pt->ps, h->lat[h->cur], h->stream);
        int rc = dispatch_err(e, "beads_collision");
        if (rc)
            return rc;
        // avedensity after a step that is a multiple of nave (main.f90:163-167),
        // taken once this bounce-back has completed the fluid nodes'
        // populations; the collision below subtracts the mean
        if (h->ave_pending) {
            rc = lbm_avedensity(h, nullptr);
            if (rc)
                return rc;
            h->ave_pending = false;
        }
        HIP_TRY(hipEventRecord(h->pev[0], h->stream));
        HIP_TRY(hipStreamWaitEvent(h->pstream, h->pev[0], 0));
        rc = dispatch_err(part_sum(pt->pg, pt->ps, h->pstream), "beads_collision sum");
        if (!rc) {
            pt->pg.fx = h->fc.fx; // beads_move's body force (pressure_force)
            pt->pg.fy = h->fc.fy;
            pt->pg.fz = h->fc.fz;
            rc = dispatch_err(part_move(pt->pg, pt->ps, h->pstream), "beads_move");
        }
        // the join point; on every return below the main stream waits for it
        // first, so no later call races the side stream's writes
        HIP_TRY(hipEventRecord(h->pev[1], h->pstream));
        if (!rc)
            rc = launch_step(h);
        if (rc) {
            (void)hipStreamWaitEvent(h->stream, h->pev[1], 0);
            return rc;
        }
        fields_from_cur(h); // macrovar
        h->rho_off = 0.0;
        const long long istep = (long long)istep0 + s + 1;
        if (nave > 0 && istep % nave == 0)
            h->ave_pending = true;
        HIP_TRY(hipStreamWaitEvent(h->stream, h->pev[1], 0));
        rc = dispatch_err(part_classify(pt->pg, pt->ps, 0, h->stream), "beads_links");
        if (!rc)
            rc = lbm_beads_filling(h);
        if (rc)
            return rc;
    }
    return check_overflow(h, "lbm_run_particles");
}

int lbm_particles_get(lbm_channel *h, double *pos, double *vel, double *omg, double *force, double *torque,
                      double *contact, double *lub)
{
    CHECK_P(h);
    const PartState &ps = h->part->ps;
    const size_t b = 3 * (size_t)h->part->pg.np * sizeof(double);
    const double *src[7] = {ps.X, ps.V, ps.W, ps.FH, ps.TQ, ps.FC, ps.FL};
    double *dst[7] = {pos, vel, omg, force, torque, contact, lub};
    for (int k = 0; k < 7; k++)
        if (dst[k])
            HIP_TRY(hipMemcpyAsync(dst[k], src[k], b, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return check_overflow(h, "lbm_particles_get");
}

int lbm_particles_links(lbm_channel *h, int *maxlink, int *counts, int *node, int *dir, double *q)
{
    CHECK_P(h);
    const PartGeom &pg = h->part->pg;
    const PartState &ps = h->part->ps;
    if (maxlink)
        *maxlink = pg.maxlink;
    const size_t nl = (size_t)pg.np * pg.maxlink;
    if (counts)
        HIP_TRY(hipMemcpyAsync(counts, ps.link_count, pg.np * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    if (node)
        HIP_TRY(hipMemcpyAsync(node, ps.link_node, nl * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    if (q)
        HIP_TRY(hipMemcpyAsync(q, ps.link_q, nl * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (dir) {
        std::vector<signed char> d(nl);
        HIP_TRY(hipMemcpyAsync(d.data(), ps.link_dir, nl, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        for (size_t k = 0; k < nl; k++)
            dir[k] = d[k];
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    return check_overflow(h, "lbm_particles_links");
}

int lbm_particles_solid(lbm_channel *h, int *solid)
{
    CHECK_P(h);
    if (!solid)
        return fail(LBM_ERR_ARG, "lbm_particles_solid: null output");
    HIP_TRY(hipMemcpyAsync(solid, h->part->ps.solid, h->g.N * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_closure_set(lbm_channel *h, const float *params, double ystar)
{
    CHECK_H(h);
    if (!params || !(ystar > 0.0))
        return fail(LBM_ERR_ARG, "lbm_closure_set: need parameters and ystar > 0");
    if (h->prec == LBM_FP64)
        return fail(LBM_ERR_ARG, "lbm_closure_set: the closure runs in FP32 arithmetic (FP32/FP16 storage)");
    if (h->part)
        return fail(LBM_ERR_STATE, "lbm_closure_set: not combined with particles");
    if (!(params[MLP_NUTMIN] > -1.0f && params[MLP_NUTMIN] <= 0.0f))
        return fail(LBM_ERR_ARG, "lbm_closure_set: the nu_t/nu floor %g must lie in (-1, 0]",
                    (double)params[MLP_NUTMIN]);
    for (int j = 0; j < MLP_H; j++)
        if (!(fabsf(params[MLP_B1 + j]) <= MLP_BIAS_MAX && fabsf(params[MLP_B2 + j]) <= MLP_BIAS_MAX))
            return fail(LBM_ERR_ARG, "lbm_closure_set: hidden bias %d beyond +-%g (or not finite)", j,
                        (double)MLP_BIAS_MAX);
    // the parameters, then the constants and the FP16 weight image of the
    // fused evaluation (mlp.h, mlp_derive)
    std::vector<float> dev(MLP_DEV_TOTAL);
    std::copy(params, params + MLP_TOTAL, dev.begin());
    mlp_derive(params, dev.data());
    if (!h->mlp)
        HIP_TRY(hipMalloc((void **)&h->mlp, MLP_DEV_TOTAL * sizeof(float)));
    HIP_TRY(hipMemcpyAsync(h->mlp, dev.data(), MLP_DEV_TOTAL * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->mlp_inv_ystar = (float)(1.0 / ystar);
    h->mlp_yscale = params[MLP_SCALE + 9];
    return LBM_OK;
}

// Closure inputs (what 0) or nu_t/nu (what 1) of every node into host memory.
static int closure_eval(lbm_channel *h, float *out, int what, const char *name)
{
    CHECK_H(h);
    if (!out)
        return fail(LBM_ERR_ARG, "%s: null output", name);
    if (!h->mlp)
        return fail(LBM_ERR_STATE, "%s: no closure set (lbm_closure_set)", name);
    if (h->ms != FROM_CUR || h->rho_off != 0.0)
        return fail(LBM_ERR_STATE, "%s: the closure needs fields from macrovar (call lbm_macrovar first)", name);
    StepArgs a;
    memset(&a, 0, sizeof(a));
    a.g = h->g;
    a.src = h->lat[h->cur];
    a.r = rates_of(h);
    a.fc = h->fc;
    a.fm = h->fm;
    a.mlp = h->mlp;
    a.inv_ystar = h->mlp_inv_ystar;
    a.yscale = h->mlp_yscale;
    const int width = what == 0 ? MLP_IN : 1;
    float *buf = nullptr;
    const uint64_t chunk = std::min<uint64_t>(COPY_NODES, h->g.N);
    HIP_TRY(hipMalloc((void **)&buf, chunk * width * sizeof(float)));
    int rc = LBM_OK;
    for (uint64_t n0 = 0; n0 < h->g.N && rc == LBM_OK; n0 += chunk) {
        const uint64_t cnt = std::min<uint64_t>(chunk, h->g.N - n0);
        rc = dispatch_err(PREC_CALL(h, closure_eval, a, n0, cnt, buf, what, h->stream), name);
        if (rc == LBM_OK) {
            hipError_t e = hipMemcpyAsync(out + n0 * width, buf, cnt * width * sizeof(float), hipMemcpyDeviceToHost,
                                          h->stream);
            if (e == hipSuccess)
                e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess)
                rc = fail(LBM_ERR_HIP, "%s: copy failed: %s", name, hipGetErrorString(e));
        }
    }
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(buf);
    return rc;
}

int lbm_closure_inputs(lbm_channel *h, float *feat) { return closure_eval(h, feat, 0, "lbm_closure_inputs"); }

int lbm_closure_nut(lbm_channel *h, float *nut) { return closure_eval(h, nut, 1, "lbm_closure_nut"); }

int lbm_closure_clear(lbm_channel *h)
{
    CHECK_H(h);
    if (h->mlp) {
        HIP_TRY(hipStreamSynchronize(h->stream));
        (void)hipFree(h->mlp);
        h->mlp = nullptr;
    }
    return LBM_OK;
}

#define CHECK_T(t)                                                   \
    do {                                                             \
        if (!(t))                                                    \
            return fail(LBM_ERR_ARG, "%s: null trainer", __func__);  \
        g_err[0] = 0;                                                \
        HIP_TRY(hipSetDevice((t)->device));                          \
    } while (0)

int lbm_trainer_create(int device, long long capacity, const float *params, lbm_trainer **out)
{
    g_err[0] = 0;
    if (!out || !params || capacity < 1 || capacity > (1ll << 31))
        return fail(LBM_ERR_ARG, "lbm_trainer_create: need params, out and 1 <= capacity <= 2^31");
    *out = nullptr;
    HIP_TRY(hipSetDevice(device));
    lbm_trainer *t = new lbm_trainer();
    memset(t, 0, sizeof(*t));
    t->device = device;
    t->capacity = capacity;
    t->max_waves = 1u << 16;
    auto cleanup = [&](int code) {
        lbm_trainer_destroy(t);
        return code;
    };
    if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(LBM_ERR_HIP, "lbm_trainer_create: stream creation failed"));
    if (hipMalloc((void **)&t->feat, (size_t)capacity * MLP_IN * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&t->target, (size_t)capacity * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&t->w, MLP_TOTAL * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&t->m, MLP_NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&t->v, MLP_NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&t->partial, (size_t)t->max_waves * (MLP_NPARAM + 1) * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&t->grad, (MLP_NPARAM + 1) * sizeof(float)) != hipSuccess)
        return cleanup(fail(LBM_ERR_NOMEM, "lbm_trainer_create: device allocation failed"));
    if (hipMemcpyAsync(t->w, params, MLP_TOTAL * sizeof(float), hipMemcpyHostToDevice, t->stream) != hipSuccess ||
        hipMemsetAsync(t->m, 0, MLP_NPARAM * sizeof(float), t->stream) != hipSuccess ||
        hipMemsetAsync(t->v, 0, MLP_NPARAM * sizeof(float), t->stream) != hipSuccess ||
        hipStreamSynchronize(t->stream) != hipSuccess)
        return cleanup(fail(LBM_ERR_HIP, "lbm_trainer_create: initialisation failed"));
    *out = t;
    return LBM_OK;
}

int lbm_trainer_destroy(lbm_trainer *t)
{
    if (!t)
        return LBM_OK;
    (void)hipSetDevice(t->device);
    if (t->stream)
        (void)hipStreamSynchronize(t->stream);
    for (void *p : {(void *)t->feat, (void *)t->target, (void *)t->w, (void *)t->m, (void *)t->v, (void *)t->partial,
                    (void *)t->grad})
        if (p)
            (void)hipFree(p);
    if (t->stream)
        (void)hipStreamDestroy(t->stream);
    delete t;
    return LBM_OK;
}

int lbm_trainer_add_filtered(lbm_trainer *t, lbm_channel *fine, double nu_coarse, double s9_coarse,
                             double ystar_coarse, long long *added)
{
    CHECK_T(t);
    if (!fine)
        return fail(LBM_ERR_ARG, "lbm_trainer_add_filtered: null channel");
    {
        HIP_TRY(hipSetDevice(fine->cfg.device));
        const int frc = flush_pending(fine);
        if (frc)
            return frc;
    }
    if (fine->cfg.device != t->device)
        return fail(LBM_ERR_ARG, "lbm_trainer_add_filtered: channel and trainer on different devices");
    if ((fine->g.lx | fine->g.ly | fine->g.lz) & 1)
        return fail(LBM_ERR_ARG, "lbm_trainer_add_filtered: fine grid %u x %u x %u must be even", fine->g.lx,
                    fine->g.ly, fine->g.lz);
    if (fine->part)
        return fail(LBM_ERR_ARG, "lbm_trainer_add_filtered: not for particle-laden channels");
    if (fine->fc.tab)
        return fail(LBM_ERR_ARG, "lbm_trainer_add_filtered: the fine channel needs a uniform force (FORCING)");
    const uint32_t lxc = fine->g.lx / 2, lyc = fine->g.ly / 2, lzc = fine->g.lz / 2;
    const long long Nc = (long long)lxc * lyc * lzc;
    if (t->count + Nc > t->capacity)
        return fail(LBM_ERR_ARG, "lbm_trainer_add_filtered: %lld samples would exceed the capacity %lld",
                    t->count + Nc, t->capacity);
    HIP_TRY(hipStreamSynchronize(fine->stream));
    // 10 filtered fields, then (per-plane targets) tau:S and S:S per node
    const bool plane = (t->target_mode == LBM_TARGET_PLANE);
    float *fields = nullptr;
    HIP_TRY(hipMalloc((void **)&fields, (size_t)Nc * (plane ? 12 : 10) * sizeof(float)));
    FilterArgs fa;
    fa.gf = fine->g;
    fa.L = fine->lat[fine->cur];
    fa.fx = fine->fc.fx;
    fa.fy = fine->fc.fy;
    fa.fz = fine->fc.fz;
    fa.lxc = lxc;
    fa.lyc = lyc;
    fa.lzc = lzc;
    fa.out = fields;
    hipError_t e = fine->prec == LBM_FP64   ? MlpLaunch<0>::filter(fa, t->stream)
                   : fine->prec == LBM_FP32 ? MlpLaunch<1>::filter(fa, t->stream)
                   : fine->prec == LBM_FP16 ? MlpLaunch<2>::filter(fa, t->stream)
                                            : MlpLaunch<3>::filter(fa, t->stream);
    FeatArgs ft;
    ft.lx = lxc;
    ft.ly = lyc;
    ft.lz = lzc;
    ft.fields = fields;
    ft.nu = (float)nu_coarse;
    ft.s9 = (float)s9_coarse;
    ft.inv_ystar = (float)(1.0 / ystar_coarse);
    std::vector<float> hp(MLP_TOTAL);
    if (e == hipSuccess)
        e = hipMemcpyAsync(hp.data(), t->w, MLP_TOTAL * sizeof(float), hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(t->stream);
    ft.yscale = hp[MLP_SCALE + 9];
    ft.scale = t->w + MLP_SCALE;
    ft.feat = t->feat;
    ft.target = t->target;
    ft.offset = (uint64_t)t->count;
    ft.ts = plane ? fields + (size_t)Nc * 10 : nullptr;
    ft.ss = plane ? fields + (size_t)Nc * 11 : nullptr;
    if (e == hipSuccess)
        e = mlp_features(ft, t->stream);
    if (e == hipSuccess && plane)
        e = mlp_plane_targets(lxc, lyc, lzc, ft.ts, ft.ss, ft.nu, t->target, ft.offset, t->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(t->stream);
    (void)hipFree(fields);
    if (e != hipSuccess)
        return fail(LBM_ERR_HIP, "lbm_trainer_add_filtered: %s", hipGetErrorString(e));
    t->count += Nc;
    if (added)
        *added = Nc;
    return LBM_OK;
}

int lbm_trainer_set_target(lbm_trainer *t, int mode)
{
    CHECK_T(t);
    if (mode != LBM_TARGET_LOCAL && mode != LBM_TARGET_PLANE)
        return fail(LBM_ERR_ARG, "lbm_trainer_set_target: mode %d is neither LBM_TARGET_LOCAL nor LBM_TARGET_PLANE",
                    mode);
    t->target_mode = mode;
    return LBM_OK;
}

int lbm_trainer_set_plane_targets(lbm_trainer *t, long long first, long long count, int nx, const float *values)
{
    CHECK_T(t);
    if (!values || nx < 1 || first < 0 || count < 0 || first + count > t->count)
        return fail(LBM_ERR_ARG, "lbm_trainer_set_plane_targets: need values, nx >= 1 and samples within the %lld held",
                    t->count);
    for (int i = 0; i < nx; i++)
        if (!std::isfinite(values[i]))
            return fail(LBM_ERR_ARG, "lbm_trainer_set_plane_targets: value %d is not finite", i);
    if (count == 0)
        return LBM_OK;
    float *dv = nullptr;
    HIP_TRY(hipMalloc((void **)&dv, (size_t)nx * sizeof(float)));
    hipError_t e = hipMemcpyAsync(dv, values, (size_t)nx * sizeof(float), hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess)
        e = mlp_fill_plane_targets(t->target, (uint64_t)first, (uint64_t)count, (uint32_t)nx, dv, t->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(t->stream);
    (void)hipFree(dv);
    if (e != hipSuccess)
        return fail(LBM_ERR_HIP, "lbm_trainer_set_plane_targets: %s", hipGetErrorString(e));
    return LBM_OK;
}

int lbm_trainer_add_samples(lbm_trainer *t, const float *features, const float *targets, long long n)
{
    CHECK_T(t);
    if (!features || !targets || n < 0 || t->count + n > t->capacity)
        return fail(LBM_ERR_ARG, "lbm_trainer_add_samples: bad arguments or capacity exceeded");
    HIP_TRY(hipMemcpyAsync(t->feat + (size_t)t->count * MLP_IN, features, (size_t)n * MLP_IN * sizeof(float),
                           hipMemcpyHostToDevice, t->stream));
    HIP_TRY(hipMemcpyAsync(t->target + t->count, targets, (size_t)n * sizeof(float), hipMemcpyHostToDevice,
                           t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    t->count += n;
    return LBM_OK;
}

int lbm_trainer_samples(lbm_trainer *t, long long *n, float *features, float *targets)
{
    CHECK_T(t);
    if (n)
        *n = t->count;
    if (features)
        HIP_TRY(hipMemcpyAsync(features, t->feat, (size_t)t->count * MLP_IN * sizeof(float), hipMemcpyDeviceToHost,
                               t->stream));
    if (targets)
        HIP_TRY(hipMemcpyAsync(targets, t->target, (size_t)t->count * sizeof(float), hipMemcpyDeviceToHost,
                               t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    return LBM_OK;
}

static int trainer_range(lbm_trainer *t, long long first, long long count, const char *fn)
{
    if (first < 0 || count < 1 || first + count > t->count)
        return fail(LBM_ERR_ARG, "%s: samples [%lld, %lld) outside [0, %lld)", fn, first, first + count, t->count);
    return LBM_OK;
}

// One gradient pass covers at most max_waves waves of 64 samples.
static int trainer_batch(lbm_trainer *t, long long batch, const char *fn)
{
    if ((batch + 63) / 64 > (long long)t->max_waves)
        return fail(LBM_ERR_ARG, "%s: batch of %lld samples exceeds %u waves of 64", fn, batch, t->max_waves);
    return LBM_OK;
}

int lbm_trainer_grad(lbm_trainer *t, long long first, long long count, float *grad, float *loss)
{
    CHECK_T(t);
    int rc = trainer_range(t, first, count, "lbm_trainer_grad");
    if (!rc)
        rc = trainer_batch(t, count, "lbm_trainer_grad");
    if (rc)
        return rc;
    HIP_TRY(mlp_grad(t->w, t->feat, t->target, (uint32_t)first, (uint32_t)count, t->partial, t->grad, t->w, t->m,
                     t->v, 1, 0.0f, 0, t->stream));
    std::vector<float> g(MLP_NPARAM + 1);
    HIP_TRY(hipMemcpyAsync(g.data(), t->grad, g.size() * sizeof(float), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    if (grad)
        memcpy(grad, g.data(), MLP_NPARAM * sizeof(float));
    if (loss)
        *loss = g[MLP_NPARAM];
    return LBM_OK;
}

int lbm_trainer_train(lbm_trainer *t, long long first, long long count, int iterations, long long batch, float lr,
                      float *loss)
{
    CHECK_T(t);
    int rc = trainer_range(t, first, count, "lbm_trainer_train");
    if (rc)
        return rc;
    if (iterations < 1 || batch < 1 || batch > count)
        return fail(LBM_ERR_ARG, "lbm_trainer_train: need iterations >= 1 and 1 <= batch <= count");
    rc = trainer_batch(t, batch, "lbm_trainer_train");
    if (rc)
        return rc;
    long long pos = 0;
    for (int it = 0; it < iterations; it++) {
        if (pos + batch > count)
            pos = 0;
        t->step++;
        HIP_TRY(mlp_grad(t->w, t->feat, t->target, (uint32_t)(first + pos), (uint32_t)batch, t->partial, t->grad,
                         t->w, t->m, t->v, t->step, lr, 1, t->stream));
        pos += batch;
    }
    float g[MLP_NPARAM + 1];
    HIP_TRY(hipMemcpyAsync(g, t->grad, sizeof(g), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    if (loss)
        *loss = g[MLP_NPARAM];
    return LBM_OK;
}

int lbm_trainer_predict(lbm_trainer *t, long long first, long long count, float *out)
{
    CHECK_T(t);
    if (!out || first < 0 || count < 1 || first + count > t->count)
        return fail(LBM_ERR_ARG, "lbm_trainer_predict: bad range or null output");
    float *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, (size_t)count * sizeof(float)));
    hipError_t e = mlp_predict(t->w, t->feat, (uint32_t)first, (uint32_t)count, d, t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, d, (size_t)count * sizeof(float), hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(t->stream);
    (void)hipFree(d);
    if (e != hipSuccess)
        return fail(LBM_ERR_HIP, "lbm_trainer_predict: %s", hipGetErrorString(e));
    return LBM_OK;
}

int lbm_trainer_params(lbm_trainer *t, float *params)
{
    CHECK_T(t);
    if (!params)
        return fail(LBM_ERR_ARG, "lbm_trainer_params: null output");
    HIP_TRY(hipMemcpyAsync(params, t->w, MLP_TOTAL * sizeof(float), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    return LBM_OK;
}

} // extern "C"
