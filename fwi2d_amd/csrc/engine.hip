// engine.hip -- host side of the C ABI declared in include/fwi2d.h.
//
// A context owns one HIP stream and every device buffer of a survey.  Shots
// run in batches of up to max_batch resident slots; each pass over a batch is
// one launch per time step covering every slot (all shots of the batch in
// one grid), and the T-step loop is captured once into a hipGraph per
// (pass kind, batch size) and replayed.  The first global shot of the batch
// reaches the kernels through a device scalar, so one captured graph serves
// every batch of the same size.
#include "fwi2d.h"
#include "kernels.h"
#include "dsp.h"
#include "layout.h"
#include "status.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {
thread_local std::string g_last_error;
}

// Staging buffer of fwi_rb_velocity_noalloc: one copy per GPU, held until the
// process ends.  In host code `scratch` names a host-side stand-in, not device
// memory: host code uses only the address hipGetSymbolAddress returns on each
// device (scratch_on_device below), never the name.  Every call on one GPU
// shares the array, so calls that overlap on one GPU race.
constexpr long kScratchFloats = 1L << 21;    // 8 MiB: padded grids up to 2,097,152 cells
static __device__ __align__(256) float scratch[kScratchFloats];

int fwi::fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

namespace {
using fwi::fail;

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(FWI_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                            \
    } while (0)

#define TRY(expr)                    \
    do {                             \
        int r_ = (expr);             \
        if (r_ != FWI_OK) return r_; \
    } while (0)

constexpr float kCflLimit = 0.5546f;   // 2-D leapfrog limit of the 8th-order stencil
using fwi::kGuard;

struct GridBuf {
    fwi::Grid g{};
    bool on = false;
    float *vpad = nullptr;       // padded base model (slot layout)
    int *src_iz = nullptr, *src_ix = nullptr;
    int *col_ptr = nullptr, *rec_iz = nullptr, *rec_id = nullptr;
    std::vector<int> h_col_ptr;  // host copy of col_ptr (staging decisions)
};

// Four buffers of one wavefield: `cur` holds the latest step, `old` the one
// before.  A single step writes over old; a launch of two or three steps
// writes its last two steps to the two free buffers.
struct Quad {
    float *b[4];
    int cur, old;
    void free_pair(int *f1, int *f2) const
    {
        int n = 0, f[2] = {0, 0};
        for (int k = 0; k < 4; ++k)
            if (k != cur && k != old) f[n++] = k;
        *f1 = f[0];
        *f2 = f[1];
    }
    void after_single() { std::swap(cur, old); }
    void after_double(int f1, int f2) { old = f1; cur = f2; }
};

enum PassKind { PASS_FWD_ABC_REC = 0, PASS_FWD_RBC = 1, PASS_FWD_RBC_REC = 2,
                PASS_REV_IMAGE = 3, PASS_REV_GRAD = 4, PASS_FWD_ABC = 5 };

}  // namespace

struct fwi_ctx {
    fwi_config cfg{};
    hipStream_t stream = nullptr;
    GridBuf grid[2];
    fwi::Layout lay{};         // buffer sizes and fill extents
    std::map<const void *, long> cap;   // floats allocated per device buffer
    float kc = 0.f;            // dt^2 / dh^2
    float idh2 = 0.f;

    float *d_vp = nullptr;     // interior model
    float *d_cabc = nullptr;   // shared coefficient, abc grid
    float *d_crb = nullptr;    // per-slot coefficient, rbc grid
    float *d_b1d = nullptr;
    float *d_wav = nullptr;
    int *d_shot0 = nullptr;
    float *d_f[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    float *d_acc_slot = nullptr, *d_ill_slot = nullptr;
    float *d_image = nullptr, *d_grad = nullptr, *d_illum = nullptr;
    float *d_obs = nullptr;    // [nsrc][nrec][nt]
    float *d_seis = nullptr;   // [max_batch][nrec][nt]
    double *d_partials = nullptr, *d_sum = nullptr;
    float *d_tmp = nullptr;    // nz*nx scratch for crops
    int npartials = 1024;

    int nsrc = 0, nrec = 0, tlag = 0;
    bool have_wavelet = false, have_geometry = false, have_model = false;
    float vmin = 0.f, vmax = 0.f;
    Quad fwd{};                // final (u_T, u_{T-1}) of the last forward pass

    // launch shapes: [0] random-boundary forward, [1] reverse, [2] sponge forward;
    // the defaults are the shapes the autotuner picks for batches of shots
    int vz[3] = {2, 2, 2}, xw[3] = {81, 81, 81};
    int nsteps[3] = {4, 4, 4};              // time steps per launch
    int use_graph = 1;
    std::map<long, hipGraphExec_t> graphs;

    float *scratch = nullptr;  // this device's copy of `scratch` (fwi_rb_velocity_noalloc)

    // FWI_RBC_GAUSSIAN with blobs: their shape, the host sampling buffers (one
    // per host thread), the table staged in pinned memory and its device copy
    fwi::RbShape rbs{};
    std::vector<fwi::RbWork> rbwork;
    std::vector<uint64_t> rb_streams;
    long blob_cap = 0;         // blobs the tables hold (max_batch streams)
    fwi::RbBlob *h_blob = nullptr, *d_blob = nullptr;
    int32_t *h_bstart = nullptr, *d_bstart = nullptr;
    hipEvent_t rb_copied = nullptr;   // the last table upload has read the staging
    bool rb_pending = false;

    bool profile = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double prof_ms[2] = {0, 0};
    int64_t prof_launches[2] = {0, 0};
    double prof_cells[2] = {0, 0};
};

namespace {

int T_of(const fwi_ctx *c) { return c->cfg.nt + c->tlag; }

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept per device.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float *g_scratch_ptr[kMaxDevices] = {};

int scratch_on_device(float **out)
{
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(FWI_ERR_HIP, "device ordinal %d outside [0, %d)", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(FWI_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<float *>(p);
    }
    *out = g_scratch_ptr[dev];
    return FWI_OK;
}

fwi::Tiling tiling_of(const fwi_ctx *c, const fwi::Grid &g, int kind)
{
    fwi::Tiling t;
    t.vz = c->vz[kind];
    t.xw = c->xw[kind];
    t.nzseg = fwi::zsegments(g.nzz, t.vz, c->nsteps[kind]);
    t.nxchunk = fwi::xchunks(g.nxx, t.vz, t.xw, c->nsteps[kind]);
    t.staged = true;
    return t;
}

// Whether every wave of a launch of `steps` steps with tiling t can stage its
// window in LDS: at most kStageCols columns and, when the launch uses the
// receiver tables, at most kStageRec receivers (kStageRec4 in the four-step
// reverse kernel) -- the windows and counts the kernels' stage_table
// computes.
bool staging_fits(const GridBuf &gb, const fwi::Tiling &t, int steps, bool tables,
                  bool reverse)
{
    if (t.vz == fwi::kTileVz) return true;         // no per-wave tables
    const int h = fwi::window_halo(steps);
    const int maxrec = (reverse && steps == 4) ? fwi::kStageRec4 : fwi::kStageRec;
    if (t.xw + 2 * h > fwi::kStageCols) return false;
    if (!tables || gb.h_col_ptr.empty()) return true;
    const int nxx = gb.g.nxx;
    for (int xc = 0; xc < t.nxchunk; ++xc) {
        const int x0 = 4 + xc * t.xw, x1 = std::min(x0 + t.xw, nxx - 4);
        const int ca = std::min(std::max(x0 - h, 0), nxx);
        const int cb = std::min(std::max(x1 + h, 0), nxx);
        if (gb.h_col_ptr[cb] - gb.h_col_ptr[ca] > maxrec) return false;
    }
    return true;
}

void drop_graphs(fwi_ctx *c)
{
    for (auto &kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
    c->graphs.clear();
}

// Device allocation zeroed on `stream` and waited for: the context stream is
// non-blocking, so a zero-fill on the legacy null stream would not be ordered
// before the kernels that use the buffer.
template <typename T>
int dalloc(T **p, size_t n, hipStream_t stream)
{
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess)
        return fail(FWI_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T),
                    hipGetErrorString(e));
    e = hipMemsetAsync(*p, 0, n * sizeof(T), stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return fail(FWI_ERR_HIP, "zero-fill failed: %s", hipGetErrorString(e));
    return FWI_OK;
}

// Zeroed device allocation of n floats owned by the context, recorded with
// its size so that every fill can be checked against it.
int cdalloc(fwi_ctx *c, float **p, long n)
{
    TRY(dalloc(p, (size_t)n, c->stream));
    c->cap[*p] = n;
    return FWI_OK;
}

// Zero-fill of the first n floats of a context buffer.  An extent beyond the
// buffer's allocation is an engine bug: it fails with FWI_ERR_INTERNAL
// instead of writing past the buffer.
int fill(fwi_ctx *c, float *p, long n)
{
    const auto it = c->cap.find(p);
    const long have = it == c->cap.end() ? -1 : it->second;
    if (n > have)
        return fail(FWI_ERR_INTERNAL, "fill of %ld floats exceeds the buffer's %ld", n, have);
    HIP_TRY(fwi::launch_fill(p, n, 0.f, c->stream));
    return FWI_OK;
}

fwi::FwdStep fwd_args(fwi_ctx *c, int kind, bool record)
{
    const GridBuf &gb = c->grid[kind];
    fwi::FwdStep a{};
    a.g = gb.g;
    a.c = (kind == FWI_BOUNDARY_ABC) ? c->d_cabc : c->d_crb;
    a.cstride = (kind == FWI_BOUNDARY_ABC) ? 0 : gb.g.fstride;
    a.b1d = c->d_b1d;
    a.abc = (kind == FWI_BOUNDARY_ABC) ? 1 : 0;
    a.src_iz = gb.src_iz;
    a.src_ix = gb.src_ix;
    a.shot0 = c->d_shot0;
    a.wav = c->d_wav;
    a.nt = c->cfg.nt;
    a.tlag = c->tlag;
    a.T = T_of(c);
    a.rec = {gb.col_ptr, gb.rec_iz, gb.rec_id, c->nrec};
    a.seis = record ? c->d_seis : nullptr;
    return a;
}

fwi::RevStep rev_args(fwi_ctx *c, int mode)
{
    const GridBuf &gb = c->grid[FWI_BOUNDARY_RBC];
    fwi::RevStep a{};
    a.g = gb.g;
    a.c = c->d_crb;
    a.cstride = gb.g.fstride;
    a.shot0 = c->d_shot0;
    a.data = (mode == 0) ? c->d_obs : c->d_seis;
    a.data_by_shot = (mode == 0) ? 1 : 0;
    a.idh2 = c->idh2;
    a.nt = c->cfg.nt;
    a.T = T_of(c);
    a.rec = {gb.col_ptr, gb.rec_iz, gb.rec_id, c->nrec};
    a.acc = c->d_acc_slot;
    a.ill = c->d_ill_slot;
    a.mode = mode;
    return a;
}

// Enqueues the launches of one pass over T steps -- the tuned number of
// steps per launch (1-4), shorter launches for the remainder -- or
// replays the captured graph of the same sequence.  `sq` is the source (or
// forward) wavefield, `rq` the receiver wavefield of a reverse pass; both
// are advanced to their final state.
int run_steps(fwi_ctx *c, int pass, int nslots, const fwi::FwdStep *fa, const fwi::RevStep *ra,
              Quad &sq, Quad *rq)
{
    const int T = T_of(c);
    const fwi::Grid &g = fa ? fa->g : ra->g;
    const int kind = fa ? (fa->abc ? 2 : 0) : 1;    // tuning slot
    const int pkind = fa ? 0 : 1;                   // profile slot: forward / reverse
    const fwi::Tiling tl = tiling_of(c, g, kind);
    const bool tables = fa ? (fa->seis != nullptr) : true;
    const int per = c->nsteps[kind];
    // the remainder of a multi-step pass runs as shorter launches, whose
    // waves own more rows each
    fwi::Tiling tls[5] = {tl, tl, tl, tl, tl};
    const GridBuf &gbuf = c->grid[kind == 2 ? FWI_BOUNDARY_ABC : FWI_BOUNDARY_RBC];
    for (int n = 1; n <= 4; ++n) {
        tls[n].nzseg = fwi::zsegments(g.nzz, tl.vz, n);
        tls[n].nxchunk = fwi::xchunks(g.nxx, tl.vz, tl.xw, n);
        tls[n].staged = staging_fits(gbuf, tls[n], n, tables, fa == nullptr);
    }
    Quad s0 = sq, r0 = rq ? *rq : sq;
    int launches = 0;
    auto enqueue = [&](bool launch) -> hipError_t {
        Quad s = s0, r = r0;
        launches = 0;
        for (int t = 0; t < T;) {
            const int n = (per == 4 && t + 4 <= T)   ? 4
                          : (per >= 3 && t + 3 <= T) ? 3
                          : (per >= 2 && t + 2 <= T) ? 2
                                                     : 1;
            int f1 = 0, f2 = 0, h1 = 0, h2 = 0;
            s.free_pair(&f1, &f2);
            r.free_pair(&h1, &h2);
            hipError_t e = hipSuccess;
            if (fa) {
                fwi::FwdStep a = *fa;
                a.cur = s.b[s.cur];
                a.old = s.b[s.old];
                a.out1 = (n >= 2) ? s.b[f1] : s.b[s.old];    // u_{t+n-1} (n >= 2)
                a.out2 = s.b[f2];                              // u_{t+n}
                if (launch) e = fwi::launch_fwd_step(a, tls[n], nslots, t, n, c->stream);
            } else {
                fwi::RevStep a = *ra;
                a.scur = s.b[s.cur];
                a.sold = s.b[s.old];
                a.sout1 = (n >= 2) ? s.b[f1] : s.b[s.old];
                a.sout2 = s.b[f2];
                a.rcur = r.b[r.cur];
                a.rold = r.b[r.old];
                a.rout1 = (n >= 2) ? r.b[h1] : r.b[r.old];
                a.rout2 = r.b[h2];
                if (launch) e = fwi::launch_rev_step(a, tls[n], nslots, t, n, c->stream);
            }
            if (e != hipSuccess) return e;
            if (n >= 2) {
                s.after_double(f1, f2);
                r.after_double(h1, h2);
            } else {
                s.after_single();
                r.after_single();
            }
            t += n;
            ++launches;
        }
        sq = s;
        if (rq) *rq = r;
        return hipSuccess;
    };
    if (c->profile) HIP_TRY(hipEventRecord(c->ev0, c->stream));
    if (c->use_graph) {
        // the captured launches bake in the starting buffers of both wavefields
        const long key = ((long)pass << 56) | ((long)nslots << 40) | ((long)per << 36) |
                         ((long)tl.vz << 32) | ((long)tl.xw << 16) | ((long)s0.cur << 12) |
                         ((long)s0.old << 8) | ((long)r0.cur << 4) | (long)r0.old;
        auto it = c->graphs.find(key);
        if (it == c->graphs.end()) {
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = enqueue(true);
            hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
            if (e != hipSuccess) return fail(FWI_ERR_HIP, "kernel launch during capture: %s",
                                             hipGetErrorString(e));
            if (e2 != hipSuccess) return fail(FWI_ERR_HIP, "hipStreamEndCapture: %s",
                                              hipGetErrorString(e2));
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return fail(FWI_ERR_HIP, "hipGraphInstantiate: %s",
                                             hipGetErrorString(e));
            it = c->graphs.emplace(key, exec).first;
        } else {
            (void)enqueue(false);      // advance the buffer bookkeeping only
        }
        HIP_TRY(hipGraphLaunch(it->second, c->stream));
    } else {
        hipError_t e = enqueue(true);
        if (e != hipSuccess) return fail(FWI_ERR_HIP, "step launch: %s", hipGetErrorString(e));
    }
    if (c->profile) {
        HIP_TRY(hipEventRecord(c->ev1, c->stream));
        HIP_TRY(hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->prof_ms[pkind] += ms;
        c->prof_launches[pkind] += launches;
        c->prof_cells[pkind] += (double)nslots * g.nzz * g.nxx * T * (pkind ? 2.0 : 1.0);
    }
    return FWI_OK;
}

// Clamp range of the random boundary: [vmin (1 - rel) - abs, vmax (1 + rel) + abs],
// the upper end capped below the CFL limit.  One rounding per operation (no
// FMA contraction) so the bounds are reproducible by a plain C statement.
void rb_bounds(const fwi_ctx *c, float *vlo, float *vhi)
{
#pragma clang fp contract(off)
    const fwi_config &k = c->cfg;
    const float rel = (k.rbc_kind == FWI_RBC_UNIFORM) ? k.rbc_varVp_rel : 0.0f;
    const float lo_scale = 1.0f - rel;
    const float hi_scale = 1.0f + rel;
    const float lo = c->vmin * lo_scale;
    const float hi = c->vmax * hi_scale;
    *vlo = lo - k.rbc_varVp;
    *vhi = hi + k.rbc_varVp;
    const float a = 0.999f * kCflLimit;
    const float b = a * k.dh;
    const float vcfl = b / k.dt;
    if (*vhi > vcfl) *vhi = vcfl;
}

int check_ready(fwi_ctx *c, bool need_obs = false)
{
    if (!c) return fail(FWI_ERR_ARG, "null context");
    if (!c->have_wavelet) return fail(FWI_ERR_STATE, "wavelet not set (fwi_set_wavelet)");
    if (!c->have_geometry) return fail(FWI_ERR_STATE, "geometry not set (fwi_set_geometry)");
    if (!c->have_model) return fail(FWI_ERR_STATE, "model not set (fwi_set_model)");
    (void)need_obs;
    return FWI_OK;
}

int check_shots(fwi_ctx *c, int shot0, int nshots)
{
    if (shot0 < 0 || nshots < 0 || shot0 + nshots > c->nsrc)
        return fail(FWI_ERR_ARG, "shot range [%d, %d) outside [0, %d)", shot0, shot0 + nshots,
                    c->nsrc);
    return FWI_OK;
}

int set_shot0(fwi_ctx *c, int shot0)
{
    HIP_TRY(fwi::launch_set_int(c->d_shot0, shot0, c->stream));
    return FWI_OK;
}

// Random-boundary velocity of slots [0, nslots) (shots shot0 + slot, which
// the device scalar d_shot0 already holds) into c_out (coefficients, per
// slot) or v_out (velocities): the uniform kind from the seed alone; the
// Gaussian kind also samples the slots' blobs on the host, stages their table
// in pinned memory and uploads it on the context stream first.
int rb_generate(fwi_ctx *c, int shot0, int nslots, uint32_t iteration, float *c_out,
                float *v_out)
{
    const GridBuf &gb = c->grid[1];
    const fwi_config &k = c->cfg;
    float vlo, vhi;
    rb_bounds(c, &vlo, &vhi);
    if (k.rbc_kind == FWI_RBC_UNIFORM) {
        HIP_TRY(fwi::launch_rb(gb.vpad, gb.g, nslots, c->d_shot0, iteration, k.rbc_seed,
                               k.rbc_varVp, k.rbc_varVp_rel, vlo, vhi, c->kc, c_out, v_out,
                               c->stream));
        return FWI_OK;
    }
    const fwi::RbShape &rs = c->rbs;
    if (rs.nbx > 0) {
        if (nslots > c->cfg.max_batch) return fail(FWI_ERR_INTERNAL, "blob table: %d slots", nslots);
        // the staging buffers are rewritten only after the previous upload read them
        if (c->rb_pending) HIP_TRY(hipEventSynchronize(c->rb_copied));
        c->rb_pending = false;
        for (int q = 0; q < nslots; ++q)
            c->rb_streams[q] = ((uint64_t)iteration << 32) | (uint32_t)(shot0 + q);
        const long n = fwi::rb_blob_table(rs, k.rbc_seed, c->rb_streams.data(), nslots,
                                          c->rbwork, c->h_blob, c->blob_cap, c->h_bstart);
        if (n < 0) return fail(FWI_ERR_INTERNAL, "blob table exceeds its %ld blobs", c->blob_cap);
        const long nstart = (long)nslots * (rs.nbx * rs.nbz + 1);
        if (n > 0)
            HIP_TRY(hipMemcpyAsync(c->d_blob, c->h_blob, n * sizeof(fwi::RbBlob),
                                   hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->d_bstart, c->h_bstart, nstart * sizeof(int32_t),
                               hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipEventRecord(c->rb_copied, c->stream));
        c->rb_pending = true;
    }
    HIP_TRY(fwi::launch_rbg(gb.vpad, gb.g, nslots, c->d_shot0, iteration, k.rbc_seed,
                            k.rbc_varVp, vlo, vhi, c->kc, k.dh, c->d_blob, c->d_bstart, rs.nbx,
                            rs.nbz, rs.bin, c_out, v_out, c->stream));
    return FWI_OK;
}

// Forward pass of one batch.  The wavefield lives in d_f[0..3]; its final
// pair (u_T, u_{T-1}) is left in c->fwd (used by the reverse pass).
int forward_batch(fwi_ctx *c, int kind, int shot0, int nslots, bool record, uint32_t iteration)
{
    TRY(set_shot0(c, shot0));
    const long nf = c->lay.pass_fill(kind, nslots);
    if (kind == FWI_BOUNDARY_RBC) TRY(rb_generate(c, shot0, nslots, iteration, c->d_crb, nullptr));
    // all four buffers start at zero in this grid's layout (rim and padding
    // cells are never written and must read as 0)
    for (int k = 0; k < 4; ++k) TRY(fill(c, c->d_f[k], nf));
    fwi::FwdStep a = fwd_args(c, kind, record);
    Quad q{{c->d_f[0], c->d_f[1], c->d_f[2], c->d_f[3]}, 1, 0};   // u_0 in b[1], u_{-1} = 0
    HIP_TRY(fwi::launch_fwd_init(a, q.b[q.cur], nslots, c->stream));
    if (record)
        HIP_TRY(fwi::launch_zero_sample0(c->d_seis, c->nrec, c->cfg.nt, nslots, c->stream));
    int pass = (kind == FWI_BOUNDARY_ABC) ? (record ? PASS_FWD_ABC_REC : PASS_FWD_ABC)
                                          : (record ? PASS_FWD_RBC_REC : PASS_FWD_RBC);
    TRY(run_steps(c, pass, nslots, &a, nullptr, q, nullptr));
    c->fwd = q;
    return FWI_OK;
}

// Reverse pass of one batch after an rbc forward pass; mode 0 image, 1 gradient.
int reverse_batch(fwi_ctx *c, int shot0, int nslots, int mode)
{
    GridBuf &gb = c->grid[FWI_BOUNDARY_RBC];
    TRY(set_shot0(c, shot0));
    const long nf = c->lay.pass_fill(FWI_BOUNDARY_RBC, nslots);
    for (int k = 4; k < 8; ++k) TRY(fill(c, c->d_f[k], nf));
    fwi::RevStep a = rev_args(c, mode);
    // source wavefield: cur = u_{T-1}, old = u_T (the forward's old / cur)
    Quad sq = c->fwd;
    std::swap(sq.cur, sq.old);
    Quad rq{{c->d_f[4], c->d_f[5], c->d_f[6], c->d_f[7]}, 1, 0};  // R_0 in b[1], R_{-1} = 0
    HIP_TRY(fwi::launch_rev_init(a, rq.b[rq.cur], nslots, c->stream));
    TRY(run_steps(c, mode == 0 ? PASS_REV_IMAGE : PASS_REV_GRAD, nslots, nullptr, &a, sq, &rq));
    float *dst = (mode == 0) ? c->d_image : c->d_grad;
    HIP_TRY(fwi::launch_reduce_slots(c->d_acc_slot, gb.g.fstride, gb.g.fstride, nslots, dst,
                                     c->stream));
    HIP_TRY(fwi::launch_reduce_slots(c->d_ill_slot, gb.g.fstride, gb.g.fstride, nslots,
                                     c->d_illum, c->stream));
    return FWI_OK;
}

// residual = obs - cal in d_seis (after an abc recorded pass); adds sum res^2.
int residual_batch(fwi_ctx *c, int shot0, int nslots, double *sum)
{
    const long per_shot = (long)c->nrec * c->cfg.nt;
    HIP_TRY(fwi::launch_residual(c->d_obs, c->d_seis, c->d_shot0, per_shot, nslots,
                                 c->d_partials, c->npartials, c->stream));
    HIP_TRY(fwi::launch_sum_partials(c->d_partials, c->npartials, c->d_sum, c->stream));
    double h = 0.0;
    HIP_TRY(hipMemcpyAsync(&h, c->d_sum, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *sum += h;
    (void)shot0;
    return FWI_OK;
}

int setup_grid(fwi_ctx *c, int kind, int nb)
{
    GridBuf &gb = c->grid[kind];
    gb.on = nb > 0;
    if (!gb.on) return FWI_OK;
    if (nb < 4) return fail(FWI_ERR_ARG, "pad width %d < 4 cells (stencil rim)", nb);
    const fwi::GridDims &d = c->lay.grid[kind];
    gb.g.nb = d.nb;
    gb.g.nzz = d.nzz;
    gb.g.nxx = d.nxx;
    gb.g.ldz = d.ldz;
    gb.g.fstride = d.fstride;
    return FWI_OK;
}

}  // namespace

extern "C" {

const char *fwi_last_error(void) { return g_last_error.c_str(); }

const char *fwi_version(void) { return "fwi2d-mi355x 0.1 (gfx950)"; }

namespace {

// Blob tables of the Gaussian boundary, sized once for max_batch streams of
// the random-boundary grid (rb_blob_bound blobs each), and the host sampling
// buffers, reserved so that sampling allocates nothing afterwards.
int setup_blobs(fwi_ctx *c)
{
    const fwi::Grid &g = c->grid[1].g;
    c->rbs = fwi::rb_shape(g.nzz, g.nxx, g.nb, c->cfg.dh, c->cfg.rbc_ratio, c->cfg.rbc_varVp);
    if (c->rbs.nbx == 0) return FWI_OK;
    const long per = fwi::rb_blob_bound(c->rbs);
    const long slots = c->cfg.max_batch;
    const long nstart = slots * ((long)c->rbs.nbx * c->rbs.nbz + 1);
    const double bytes = (double)per * slots * sizeof(fwi::RbBlob) + (double)nstart * 4;
    if (bytes > 512.0 * 1024 * 1024 || (double)per * slots > INT32_MAX)
        return fail(FWI_ERR_ARG, "rbc_ratio %g m is too small for this grid (%ld blobs per shot)",
                    c->cfg.rbc_ratio, per);
    c->blob_cap = per * slots;
    const unsigned hw = std::thread::hardware_concurrency();
    const int threads = (int)std::max(1u, std::min(16u, hw));
    c->rbwork.resize(threads);
    for (auto &w : c->rbwork) w.reserve(c->rbs);
    c->rb_streams.resize(slots);
    HIP_TRY(hipMalloc(&c->d_blob, c->blob_cap * sizeof(fwi::RbBlob)));
    HIP_TRY(hipMalloc(&c->d_bstart, nstart * sizeof(int32_t)));
    HIP_TRY(hipHostMalloc(&c->h_blob, c->blob_cap * sizeof(fwi::RbBlob), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(&c->h_bstart, nstart * sizeof(int32_t), hipHostMallocDefault));
    HIP_TRY(hipEventCreateWithFlags(&c->rb_copied, hipEventDisableTiming));
    return FWI_OK;
}

}  // namespace

int fwi_create(const fwi_config *cfg, fwi_ctx **out)
{
    if (!cfg || !out) return fail(FWI_ERR_ARG, "null argument");
    *out = nullptr;
    if (cfg->nz < 1 || cfg->nx < 1 || !(cfg->dh > 0.f) || cfg->nt < 2 || !(cfg->dt > 0.f))
        return fail(FWI_ERR_ARG, "bad grid: nz=%d nx=%d dh=%g nt=%d dt=%g", cfg->nz, cfg->nx,
                    cfg->dh, cfg->nt, cfg->dt);
    if (cfg->max_batch < 1) return fail(FWI_ERR_ARG, "max_batch must be >= 1");
    if (cfg->abc_nb <= 0 && cfg->rbc_nb <= 0)
        return fail(FWI_ERR_ARG, "neither an abc nor an rbc grid is configured");
    if (cfg->rbc_kind != FWI_RBC_UNIFORM && cfg->rbc_kind != FWI_RBC_GAUSSIAN)
        return fail(FWI_ERR_ARG, "unknown rbc_kind %d", cfg->rbc_kind);
    if (cfg->rbc_kind == FWI_RBC_GAUSSIAN && (!(cfg->rbc_ratio >= 0.f) || !(cfg->rbc_varVp >= 0.f)))
        return fail(FWI_ERR_ARG, "rbc_ratio and rbc_varVp must be >= 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(FWI_ERR_HIP, "no HIP device available");

    fwi_ctx *c = new fwi_ctx();
    c->cfg = *cfg;
    c->kc = (float)((double)cfg->dt * cfg->dt / ((double)cfg->dh * cfg->dh));
    c->idh2 = (float)(1.0 / ((double)cfg->dh * cfg->dh));
    c->lay = fwi::make_layout(*cfg);
    int r = setup_grid(c, FWI_BOUNDARY_ABC, cfg->abc_nb);
    if (r == FWI_OK) r = setup_grid(c, FWI_BOUNDARY_RBC, cfg->rbc_nb);
    if (r != FWI_OK) { delete c; return r; }
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return fail(FWI_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
    (void)hipEventCreate(&c->ev0);
    (void)hipEventCreate(&c->ev1);

    const long n_int = (long)cfg->nz * cfg->nx;
    r = FWI_OK;
    if (r == FWI_OK) r = dalloc(&c->d_vp, n_int, c->stream);
    if (r == FWI_OK) r = dalloc(&c->d_tmp, n_int, c->stream);
    if (r == FWI_OK) r = dalloc(&c->d_shot0, 1, c->stream);
    if (r == FWI_OK) r = dalloc(&c->d_wav, cfg->nt, c->stream);
    // every buffer a step kernel reads carries kGuard zeroed floats after its
    // last slot: the lanes of a wave past the last row of the last column
    // load there (the loads are not masked per lane)
    for (int k = 0; k < 8 && r == FWI_OK; ++k) r = cdalloc(c, &c->d_f[k], c->lay.wavefield());
    for (int kind = 0; kind < 2 && r == FWI_OK; ++kind)
        if (c->grid[kind].on) r = dalloc(&c->grid[kind].vpad, c->grid[kind].g.fstride, c->stream);
    if (r == FWI_OK && c->grid[0].on) {
        std::vector<float> b1d;
        fwi::cerjan_table(cfg->abc_nb, cfg->abc_factor, b1d);
        r = dalloc(&c->d_b1d, b1d.size(), c->stream);
        if (r == FWI_OK)
            r = hipMemcpy(c->d_b1d, b1d.data(), b1d.size() * 4, hipMemcpyHostToDevice) == hipSuccess
                    ? FWI_OK : fail(FWI_ERR_HIP, "b1d upload");
        if (r == FWI_OK) r = cdalloc(c, &c->d_cabc, c->lay.abc_coef());
    }
    if (r == FWI_OK && c->grid[1].on) {
        r = cdalloc(c, &c->d_crb, c->lay.rb_coef());
        if (r == FWI_OK) r = cdalloc(c, &c->d_acc_slot, c->lay.slot_acc());
        if (r == FWI_OK) r = cdalloc(c, &c->d_ill_slot, c->lay.slot_acc());
        if (r == FWI_OK) r = cdalloc(c, &c->d_image, c->lay.shot_sum());
        if (r == FWI_OK) r = cdalloc(c, &c->d_grad, c->lay.shot_sum());
        if (r == FWI_OK) r = cdalloc(c, &c->d_illum, c->lay.shot_sum());
    }
    if (r == FWI_OK) r = dalloc(&c->d_partials, c->npartials, c->stream);
    if (r == FWI_OK) r = dalloc(&c->d_sum, 1, c->stream);
    if (r == FWI_OK) r = scratch_on_device(&c->scratch);
    if (r == FWI_OK && c->grid[1].on && cfg->rbc_kind == FWI_RBC_GAUSSIAN)
        r = setup_blobs(c);
    if (r != FWI_OK) { fwi_destroy(c); return r; }
    // default launch shapes (the autotuner refines them, see the fields)
    *out = c;
    return FWI_OK;
}

int fwi_destroy(fwi_ctx *c)
{
    if (!c) return FWI_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    void *bufs[] = {c->d_vp, c->d_cabc, c->d_crb, c->d_b1d, c->d_wav, c->d_shot0, c->d_f[0],
                    c->d_f[1], c->d_f[2], c->d_f[3], c->d_f[4], c->d_f[5], c->d_f[6], c->d_f[7],
                    c->d_acc_slot, c->d_ill_slot, c->d_image,
                    c->d_grad, c->d_illum, c->d_obs, c->d_seis, c->d_partials, c->d_sum, c->d_tmp};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    for (auto &gb : c->grid) {
        void *gp[] = {gb.vpad, gb.src_iz, gb.src_ix, gb.col_ptr, gb.rec_iz, gb.rec_id};
        for (void *p : gp)
            if (p) (void)hipFree(p);
    }
    if (c->d_blob) (void)hipFree(c->d_blob);
    if (c->d_bstart) (void)hipFree(c->d_bstart);
    if (c->h_blob) (void)hipHostFree(c->h_blob);
    if (c->h_bstart) (void)hipHostFree(c->h_bstart);
    if (c->rb_copied) (void)hipEventDestroy(c->rb_copied);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return FWI_OK;
}

int fwi_ricker_wavelet(int32_t nt, float dt, float fmax, float *w_out, int32_t *tlag_out)
{
    if (nt < 2 || !(dt > 0.f) || !(fmax > 0.f) || !w_out)
        return fail(FWI_ERR_ARG, "bad wavelet arguments");
    std::vector<double> w;
    const int tlag = fwi::ricker_half_derivative(nt, dt, fmax, w);
    for (int k = 0; k < nt; ++k) w_out[k] = (float)w[k];
    if (tlag_out) *tlag_out = tlag;
    return FWI_OK;
}

int fwi_set_wavelet(fwi_ctx *c, const float *w, int32_t tlag)
{
    if (!c || !w) return fail(FWI_ERR_ARG, "null argument");
    if (tlag < 1) return fail(FWI_ERR_ARG, "tlag must be >= 1");
    std::vector<float> h(c->cfg.nt);
    for (int k = 0; k < c->cfg.nt; ++k) h[k] = c->idh2 * w[k];   // idh2 * w[t] (modeling.cu:317)
    HIP_TRY(hipMemcpy(c->d_wav, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    if (tlag != c->tlag) drop_graphs(c);
    c->tlag = tlag;
    c->have_wavelet = true;
    return FWI_OK;
}

int fwi_set_geometry(fwi_ctx *c, int32_t nsrc, const float *xsrc, const float *zsrc,
                     int32_t nrec, const float *xrec, const float *zrec)
{
    if (!c || nsrc < 1 || nrec < 1 || !xsrc || !zsrc || !xrec || !zrec)
        return fail(FWI_ERR_ARG, "bad geometry arguments");
    const float dh = c->cfg.dh;
    // Grid indices and receiver tables of every configured grid, validated
    // before anything in the context changes: a rejected geometry leaves the
    // previous one in place.
    struct Tables {
        std::vector<int> siz, six, col_ptr, s_iz, s_id;
    } tab[2];
    for (int kind = 0; kind < 2; ++kind) {
        GridBuf &gb = c->grid[kind];
        if (!gb.on) continue;
        Tables &t = tab[kind];
        const int nb = gb.g.nb;
        auto index = [&](float x) { return (int)((x + 0.5f * dh) / dh) + nb; };
        t.siz.resize(nsrc);
        t.six.resize(nsrc);
        for (int s = 0; s < nsrc; ++s) {
            t.siz[s] = index(zsrc[s]);
            t.six[s] = index(xsrc[s]);
            if (t.siz[s] < 4 || t.siz[s] >= gb.g.nzz - 4 || t.six[s] < 4 || t.six[s] >= gb.g.nxx - 4)
                return fail(FWI_ERR_ARG, "source %d at (x=%g, z=%g) m falls outside the grid",
                            s, xsrc[s], zsrc[s]);
        }
        std::vector<int> riz(nrec), rix(nrec), order(nrec);
        for (int r = 0; r < nrec; ++r) {
            riz[r] = index(zrec[r]);
            rix[r] = index(xrec[r]);
            if (riz[r] < 4 || riz[r] >= gb.g.nzz - 4 || rix[r] < 4 || rix[r] >= gb.g.nxx - 4)
                return fail(FWI_ERR_ARG, "receiver %d at (x=%g, z=%g) m falls outside the grid",
                            r, xrec[r], zrec[r]);
            order[r] = r;
        }
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return rix[a] != rix[b] ? rix[a] < rix[b] : riz[a] < riz[b];
        });
        t.col_ptr.assign(gb.g.nxx + 1, 0);
        t.s_iz.resize(nrec);
        t.s_id.resize(nrec);
        for (int r = 0; r < nrec; ++r) t.col_ptr[rix[r] + 1]++;
        for (int j = 0; j < gb.g.nxx; ++j) t.col_ptr[j + 1] += t.col_ptr[j];
        for (int k = 0; k < nrec; ++k) {
            t.s_iz[k] = riz[order[k]];
            t.s_id[k] = order[k];
        }
    }
    // From here on the context holds no usable geometry until every buffer
    // is in place (a failed allocation leaves it unset, not half-updated).
    c->have_geometry = false;
    drop_graphs(c);
    for (int kind = 0; kind < 2; ++kind) {
        GridBuf &gb = c->grid[kind];
        if (!gb.on) continue;
        const Tables &t = tab[kind];
        void *old[] = {gb.src_iz, gb.src_ix, gb.col_ptr, gb.rec_iz, gb.rec_id};
        for (void *p : old)
            if (p) (void)hipFree(p);
        gb.src_iz = gb.src_ix = gb.col_ptr = gb.rec_iz = gb.rec_id = nullptr;
        gb.h_col_ptr.clear();
        TRY(dalloc(&gb.src_iz, nsrc, c->stream));
        TRY(dalloc(&gb.src_ix, nsrc, c->stream));
        TRY(dalloc(&gb.col_ptr, gb.g.nxx + 1, c->stream));
        TRY(dalloc(&gb.rec_iz, nrec, c->stream));
        TRY(dalloc(&gb.rec_id, nrec, c->stream));
        HIP_TRY(hipMemcpy(gb.src_iz, t.siz.data(), nsrc * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(gb.src_ix, t.six.data(), nsrc * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(gb.col_ptr, t.col_ptr.data(), t.col_ptr.size() * 4,
                          hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(gb.rec_iz, t.s_iz.data(), nrec * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(gb.rec_id, t.s_id.data(), nrec * 4, hipMemcpyHostToDevice));
        gb.h_col_ptr = t.col_ptr;
    }
    if (c->d_obs) { (void)hipFree(c->d_obs); c->d_obs = nullptr; }
    if (c->d_seis) { (void)hipFree(c->d_seis); c->d_seis = nullptr; }
    const long per_shot = (long)nrec * c->cfg.nt;
    TRY(dalloc(&c->d_obs, (long)nsrc * per_shot, c->stream));
    TRY(dalloc(&c->d_seis, (long)c->cfg.max_batch * per_shot, c->stream));
    c->nsrc = nsrc;
    c->nrec = nrec;
    c->have_geometry = true;
    return FWI_OK;
}

int fwi_set_model(fwi_ctx *c, const float *vp)
{
    if (!c || !vp) return fail(FWI_ERR_ARG, "null argument");
    const long n = (long)c->cfg.nz * c->cfg.nx;
    float vmin = 1e30f, vmax = -1e30f;
    for (long k = 0; k < n; ++k) {
        if (!(vp[k] > 0.f) || !std::isfinite(vp[k]))
            return fail(FWI_ERR_NUMERIC, "velocity %g at cell %ld is not positive and finite",
                        vp[k], k);
        vmin = std::min(vmin, vp[k]);
        vmax = std::max(vmax, vp[k]);
    }
    const float courant = vmax * c->cfg.dt / c->cfg.dh;
    if (courant > kCflLimit)
        return fail(FWI_ERR_NUMERIC, "CFL violated: vmax*dt/dh = %g > %g", courant, kCflLimit);
    HIP_TRY(hipMemcpyAsync(c->d_vp, vp, n * 4, hipMemcpyHostToDevice, c->stream));
    for (int kind = 0; kind < 2; ++kind) {
        GridBuf &gb = c->grid[kind];
        if (!gb.on) continue;
        HIP_TRY(fwi::launch_pad_model(c->d_vp, c->cfg.nz, c->cfg.nx, gb.g, gb.vpad, c->stream));
    }
    if (c->grid[0].on)
        HIP_TRY(fwi::launch_coef(c->grid[0].vpad, c->grid[0].g, c->kc, c->d_cabc, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->vmin = vmin;
    c->vmax = vmax;
    c->have_model = true;
    return FWI_OK;
}

int fwi_set_observed(fwi_ctx *c, int32_t shot0, int32_t nshots, const float *data)
{
    if (!c || !data) return fail(FWI_ERR_ARG, "null argument");
    if (!c->have_geometry) return fail(FWI_ERR_STATE, "geometry not set (fwi_set_geometry)");
    TRY(check_shots(c, shot0, nshots));
    const long per_shot = (long)c->nrec * c->cfg.nt;
    HIP_TRY(hipMemcpy(c->d_obs + (long)shot0 * per_shot, data, (size_t)nshots * per_shot * 4,
                      hipMemcpyHostToDevice));
    return FWI_OK;
}

int fwi_modeling(fwi_ctx *c, int32_t boundary, int32_t shot0, int32_t nshots, float *seis_out)
{
    TRY(check_ready(c));
    TRY(check_shots(c, shot0, nshots));
    if (boundary != FWI_BOUNDARY_ABC && boundary != FWI_BOUNDARY_RBC)
        return fail(FWI_ERR_ARG, "unknown boundary %d", boundary);
    if (!c->grid[boundary].on) return fail(FWI_ERR_ARG, "boundary %d grid not configured", boundary);
    const long per_shot = (long)c->nrec * c->cfg.nt;
    for (int b0 = shot0; b0 < shot0 + nshots; b0 += c->cfg.max_batch) {
        const int ns = std::min(c->cfg.max_batch, shot0 + nshots - b0);
        TRY(forward_batch(c, boundary, b0, ns, seis_out != nullptr, 0));
        if (seis_out)
            HIP_TRY(hipMemcpyAsync(seis_out + (long)(b0 - shot0) * per_shot, c->d_seis,
                                   (size_t)ns * per_shot * 4, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FWI_OK;
}

int fwi_migration(fwi_ctx *c, int32_t shot0, int32_t nshots)
{
    TRY(check_ready(c));
    TRY(check_shots(c, shot0, nshots));
    if (!c->grid[1].on) return fail(FWI_ERR_ARG, "random-boundary grid not configured");
    for (int b0 = shot0; b0 < shot0 + nshots; b0 += c->cfg.max_batch) {
        const int ns = std::min(c->cfg.max_batch, shot0 + nshots - b0);
        TRY(forward_batch(c, FWI_BOUNDARY_RBC, b0, ns, false, 0));
        TRY(reverse_batch(c, b0, ns, 0));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FWI_OK;
}

int fwi_gradient(fwi_ctx *c, int32_t shot0, int32_t nshots, uint32_t iteration, double *sum_res2)
{
    TRY(check_ready(c));
    TRY(check_shots(c, shot0, nshots));
    if (!c->grid[0].on || !c->grid[1].on)
        return fail(FWI_ERR_ARG, "the gradient needs both the abc and the rbc grid");
    double acc = 0.0;
    for (int b0 = shot0; b0 < shot0 + nshots; b0 += c->cfg.max_batch) {
        const int ns = std::min(c->cfg.max_batch, shot0 + nshots - b0);
        TRY(forward_batch(c, FWI_BOUNDARY_ABC, b0, ns, true, 0));
        TRY(residual_batch(c, b0, ns, &acc));
        TRY(forward_batch(c, FWI_BOUNDARY_RBC, b0, ns, false, iteration));
        TRY(reverse_batch(c, b0, ns, 1));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (sum_res2) *sum_res2 += acc;
    return FWI_OK;
}

int fwi_misfit(fwi_ctx *c, int32_t shot0, int32_t nshots, double *sum_res2)
{
    TRY(check_ready(c));
    TRY(check_shots(c, shot0, nshots));
    if (!c->grid[0].on) return fail(FWI_ERR_ARG, "the misfit needs the abc grid");
    double acc = 0.0;
    for (int b0 = shot0; b0 < shot0 + nshots; b0 += c->cfg.max_batch) {
        const int ns = std::min(c->cfg.max_batch, shot0 + nshots - b0);
        TRY(forward_batch(c, FWI_BOUNDARY_ABC, b0, ns, true, 0));
        TRY(residual_batch(c, b0, ns, &acc));
    }
    if (sum_res2) *sum_res2 += acc;
    return FWI_OK;
}

int fwi_reset(fwi_ctx *c, int32_t what)
{
    if (!c) return fail(FWI_ERR_ARG, "null context");
    if (!c->grid[1].on) return FWI_OK;
    const long fs = c->lay.shot_sum();
    if (what & FWI_RESET_IMAGE) TRY(fill(c, c->d_image, fs));
    if (what & FWI_RESET_GRADIENT) TRY(fill(c, c->d_grad, fs));
    if (what & FWI_RESET_ILLUM) TRY(fill(c, c->d_illum, fs));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FWI_OK;
}

static int crop_to_host(fwi_ctx *c, const float *src, float *dst)
{
    const long n = (long)c->cfg.nz * c->cfg.nx;
    HIP_TRY(fwi::launch_crop(src, c->grid[1].g, c->cfg.nz, c->cfg.nx, c->d_tmp, c->stream));
    HIP_TRY(hipMemcpyAsync(dst, c->d_tmp, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FWI_OK;
}

int fwi_get_image(fwi_ctx *c, float *image_out, float *raw_image, float *raw_illum)
{
    if (!c) return fail(FWI_ERR_ARG, "null context");
    if (!c->grid[1].on) return fail(FWI_ERR_ARG, "random-boundary grid not configured");
    const int nz = c->cfg.nz, nx = c->cfg.nx;
    const long n = (long)nz * nx;
    std::vector<float> im(n), il(n);
    TRY(crop_to_host(c, c->d_image, im.data()));
    TRY(crop_to_host(c, c->d_illum, il.data()));
    if (raw_image) std::memcpy(raw_image, im.data(), n * 4);
    if (raw_illum) std::memcpy(raw_illum, il.data(), n * 4);
    if (image_out) {
        std::vector<double> out;
        fwi::second_derivative_z(im.data(), il.data(), nz, nx, c->cfg.dh, 0, out);
        for (long k = 0; k < n; ++k) image_out[k] = (float)out[k];
    }
    return FWI_OK;
}

int fwi_get_gradient(fwi_ctx *c, float zmask, float *grad_out, float *raw_grad,
                     float *raw_illum, double *gmax_out)
{
    if (!c) return fail(FWI_ERR_ARG, "null context");
    if (!c->grid[1].on) return fail(FWI_ERR_ARG, "random-boundary grid not configured");
    const int nz = c->cfg.nz, nx = c->cfg.nx;
    const long n = (long)nz * nx;
    std::vector<float> gr(n), il(n);
    TRY(crop_to_host(c, c->d_grad, gr.data()));
    TRY(crop_to_host(c, c->d_illum, il.data()));
    // the kernels accumulate R * c * lap(S) = dt * (dt R v^2 lap(S)); undo one dt
    const float inv_dt = 1.0f / c->cfg.dt;
    for (long k = 0; k < n; ++k) gr[k] *= inv_dt;
    if (raw_grad) std::memcpy(raw_grad, gr.data(), n * 4);
    if (raw_illum) std::memcpy(raw_illum, il.data(), n * 4);
    std::vector<double> out;
    const int izm = (int)(zmask / c->cfg.dh);
    fwi::second_derivative_z(gr.data(), il.data(), nz, nx, c->cfg.dh, izm, out);
    double gmax = 0.0;
    for (long k = 0; k < n; ++k) gmax = std::max(gmax, std::fabs(out[k]));
    if (gmax_out) *gmax_out = gmax;
    if (grad_out) {
        const double s = gmax > 0.0 ? 1.0 / gmax : 0.0;
        for (long k = 0; k < n; ++k) grad_out[k] = (float)(out[k] * s);
    }
    return FWI_OK;
}

int fwi_parabola_step(const double *alpha, const double *f, double *step)
{
    if (!alpha || !f || !step) return fail(FWI_ERR_ARG, "null argument");
    if (!fwi::parabola_vertex(alpha, f, step))
        return fail(FWI_ERR_NUMERIC, "degenerate parabola (collinear or non-finite misfits)");
    return FWI_OK;
}

static int accumulator(fwi_ctx *c, int32_t what, float **dev)
{
    if (!c) return fail(FWI_ERR_ARG, "null context");
    if (!c->grid[1].on) return fail(FWI_ERR_ARG, "random-boundary grid not configured");
    if (what == FWI_RESET_IMAGE) *dev = c->d_image;
    else if (what == FWI_RESET_GRADIENT) *dev = c->d_grad;
    else if (what == FWI_RESET_ILLUM) *dev = c->d_illum;
    else return fail(FWI_ERR_ARG, "what must be one of image, gradient, illumination");
    return FWI_OK;
}

int fwi_get_accumulator(fwi_ctx *c, int32_t what, float *out)
{
    float *dev = nullptr;
    TRY(accumulator(c, what, &dev));
    if (!out) return fail(FWI_ERR_ARG, "null output");
    const fwi::Grid &g = c->grid[1].g;
    HIP_TRY(hipMemcpy2DAsync(out, (size_t)g.nzz * 4, dev, (size_t)g.ldz * 4, (size_t)g.nzz * 4,
                             g.nxx, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FWI_OK;
}

int fwi_set_accumulator(fwi_ctx *c, int32_t what, const float *in)
{
    float *dev = nullptr;
    TRY(accumulator(c, what, &dev));
    if (!in) return fail(FWI_ERR_ARG, "null input");
    const fwi::Grid &g = c->grid[1].g;
    HIP_TRY(hipMemcpy2DAsync(dev, (size_t)g.ldz * 4, in, (size_t)g.nzz * 4, (size_t)g.nzz * 4,
                             g.nxx, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FWI_OK;
}

int fwi_guarded_step(const double *alpha, const double *f, const float *m, const float *g,
                     int64_t n, double m_min, double *step, int32_t *rule)
{
    if (!alpha || !f || !m || !g || !step || !rule || n < 1 || !(m_min >= 0.0))
        return fail(FWI_ERR_ARG, "bad arguments");
    if (!std::isfinite(f[0])) return fail(FWI_ERR_NUMERIC, "non-finite misfit at alpha_0");
    for (int k = 1; k < 3; ++k)
        if (std::isnan(f[k]) || f[k] < 0.0) return fail(FWI_ERR_NUMERIC, "invalid trial misfit");
    if (!(alpha[0] < alpha[1] && alpha[1] < alpha[2]))
        return fail(FWI_ERR_ARG, "trial steps must increase");
    const float lo = (float)m_min;
    auto admissible = [&](double s) {
        for (int64_t k = 0; k < n; ++k) {
            const float mk = m[k] - (float)s * g[k];
            if (!(mk > 0.f && mk >= lo)) return false;
        }
        return true;
    };
    const bool finite = std::isfinite(f[1]) && std::isfinite(f[2]);
    if (finite) {
        // curvature of the parabola through the three points (> 0: convex)
        const double s01 = (f[1] - f[0]) / (alpha[1] - alpha[0]);
        const double s02 = (f[2] - f[0]) / (alpha[2] - alpha[0]);
        const double curv = (s02 - s01) / (alpha[2] - alpha[1]);
        double v = 0.0;
        if (curv > 0.0 && fwi::parabola_vertex(alpha, f, &v) && v > 0.0 && v <= 4.0 * alpha[2] &&
            admissible(v)) {
            *step = v;
            *rule = 0;
            return FWI_OK;
        }
    }
    const int order[2] = {(f[2] < f[1]) ? 2 : 1, (f[2] < f[1]) ? 1 : 2};
    for (int kb : order)
        if (f[kb] < f[0] && admissible(alpha[kb])) {
            *step = alpha[kb];
            *rule = 1;
            return FWI_OK;
        }
    *step = 0.0;
    *rule = 2;
    return FWI_OK;
}

}  // extern "C"

namespace {

// Random-boundary velocity of one shot written to the device buffer d_out
// (nzz * nxx floats) and copied to the host array out.
int rb_velocity_via(fwi_ctx *c, int32_t shot, uint32_t iteration, float *d_out, float *out)
{
    const GridBuf &gb = c->grid[1];
    const long ncell = (long)gb.g.nzz * gb.g.nxx;
    TRY(set_shot0(c, shot));
    TRY(rb_generate(c, shot, 1, iteration, nullptr, d_out));
    hipError_t e = hipMemcpyAsync(out, d_out, ncell * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(FWI_ERR_HIP, "rb velocity: %s", hipGetErrorString(e));
    return FWI_OK;
}

int check_rb_velocity(fwi_ctx *c, int32_t shot, const float *out)
{
    TRY(check_ready(c));
    TRY(check_shots(c, shot, 1));
    if (!c->grid[1].on) return fail(FWI_ERR_ARG, "random-boundary grid not configured");
    if (!out) return fail(FWI_ERR_ARG, "null output");
    return FWI_OK;
}

}  // namespace

extern "C" {

int fwi_rb_blobs(const fwi_config *cfg, int32_t shot, uint32_t iteration, float *out,
                 int64_t cap, int64_t *count)
{
    if (!cfg || !count || (cap > 0 && !out)) return fail(FWI_ERR_ARG, "null argument");
    if (cfg->nz < 1 || cfg->nx < 1 || !(cfg->dh > 0.f) || cfg->rbc_nb <= 0 || cap < 0)
        return fail(FWI_ERR_ARG, "bad grid or capacity");
    const fwi::GridDims d = fwi::grid_dims(cfg->nz, cfg->nx, cfg->rbc_nb);
    const fwi::RbShape rs =
        fwi::rb_shape(d.nzz, d.nxx, d.nb, cfg->dh, cfg->rbc_ratio, cfg->rbc_varVp);
    fwi::RbWork w;
    w.reserve(rs);
    fwi::rb_blobs(rs, cfg->rbc_seed, ((uint64_t)iteration << 32) | (uint32_t)shot, w);
    *count = (int64_t)w.blobs.size();
    for (int64_t k = 0; k < *count && k < cap; ++k) {
        const fwi::RbBlob &b = w.blobs[k];
        float *o = out + 5 * k;
        o[0] = b.x; o[1] = b.z; o[2] = b.c2; o[3] = b.h; o[4] = b.a;
    }
    return FWI_OK;
}

int fwi_rb_velocity(fwi_ctx *c, int32_t shot, uint32_t iteration, float *out)
{
    TRY(check_rb_velocity(c, shot, out));
    const long ncell = (long)c->grid[1].g.nzz * c->grid[1].g.nxx;
    float *d_out = nullptr;
    TRY(dalloc(&d_out, ncell, c->stream));
    const int r = rb_velocity_via(c, shot, iteration, d_out, out);
    (void)hipFree(d_out);
    return r;
}

int fwi_rb_velocity_noalloc(fwi_ctx *c, int32_t shot, uint32_t iteration, float *out)
{
    TRY(check_rb_velocity(c, shot, out));
    const long ncell = (long)c->grid[1].g.nzz * c->grid[1].g.nxx;
    if (ncell > kScratchFloats)
        return fail(FWI_ERR_ARG,
                    "random-boundary grid of %ld cells exceeds the %ld-cell staging buffer of "
                    "fwi_rb_velocity_noalloc (use fwi_rb_velocity)", ncell, kScratchFloats);
    if (!c->scratch) return fail(FWI_ERR_STATE, "no staging buffer on this context's device");
    // the kernel writes every one of the ncell values: nothing to clear first
    return rb_velocity_via(c, shot, iteration, c->scratch, out);
}

int fwi_get_config(fwi_ctx *c, fwi_config *out)
{
    if (!c || !out) return fail(FWI_ERR_ARG, "null argument");
    *out = c->cfg;
    return FWI_OK;
}

int fwi_geometry_size(fwi_ctx *c, int32_t *nsrc, int32_t *nrec)
{
    if (!c) return fail(FWI_ERR_ARG, "null context");
    if (nsrc) *nsrc = c->have_geometry ? c->nsrc : 0;
    if (nrec) *nrec = c->have_geometry ? c->nrec : 0;
    return FWI_OK;
}

int fwi_grid_shape(fwi_ctx *c, int32_t boundary, int32_t *nzz, int32_t *nxx, int32_t *nb)
{
    if (!c || boundary < 0 || boundary > 1) return fail(FWI_ERR_ARG, "bad argument");
    const GridBuf &gb = c->grid[boundary];
    if (nzz) *nzz = gb.on ? gb.g.nzz : 0;
    if (nxx) *nxx = gb.on ? gb.g.nxx : 0;
    if (nb) *nb = gb.on ? gb.g.nb : 0;
    return FWI_OK;
}

int fwi_synchronize(fwi_ctx *c)
{
    if (!c) return fail(FWI_ERR_ARG, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FWI_OK;
}

int fwi_set_tuning(fwi_ctx *c, int32_t kernel, int32_t steps, int32_t vz, int32_t xw,
                   int32_t use_graph)
{
    if (!c) return fail(FWI_ERR_ARG, "null context");
    if (kernel < -1 || kernel > 2) return fail(FWI_ERR_ARG, "kernel must be -1, 0, 1 or 2");
    if (vz > 0 && vz != 1 && vz != 2 && vz != 4 && vz != fwi::kTileVz)
        return fail(FWI_ERR_ARG, "vz must be 1, 2, 4 or %d (forward LDS tiles)", fwi::kTileVz);
    if (steps > 4) return fail(FWI_ERR_ARG, "steps per launch must be 1, 2, 3 or 4");
    int ns[3], vs[3], ws[3];
    for (int k = 0; k < 3; ++k) {
        const bool sel = kernel == -1 || kernel == k;
        ns[k] = (sel && steps > 0) ? steps : c->nsteps[k];
        vs[k] = (sel && vz > 0) ? vz : c->vz[k];
        ws[k] = (sel && xw > 0) ? xw : c->xw[k];
    }
    if (ns[1] >= 3 && vs[1] == 4)
        return fail(FWI_ERR_ARG, "the three- and four-step reverse kernels take 1 or 2 rows per lane");
    for (int k = 0; k < 3; ++k)
        if (vs[k] == fwi::kTileVz && ws[k] != 4 && ws[k] != 8 && ws[k] != 16)
            return fail(FWI_ERR_ARG, "with vz %d, xw is the tile columns per wave: 4, 8 or 16",
                        fwi::kTileVz);
    for (int k = 0; k < 3; ++k) {
        c->nsteps[k] = ns[k];
        c->vz[k] = vs[k];
        c->xw[k] = ws[k];
    }
    if (use_graph >= 0) c->use_graph = use_graph ? 1 : 0;
    drop_graphs(c);
    return FWI_OK;
}

int fwi_get_tuning(fwi_ctx *c, int32_t kernel, int32_t *steps, int32_t *vz, int32_t *xw,
                   int32_t *use_graph)
{
    if (!c) return fail(FWI_ERR_ARG, "null context");
    if (kernel < 0 || kernel > 2) return fail(FWI_ERR_ARG, "kernel must be 0, 1 or 2");
    if (steps) *steps = c->nsteps[kernel];
    if (vz) *vz = c->vz[kernel];
    if (xw) *xw = c->xw[kernel];
    if (use_graph) *use_graph = c->use_graph;
    return FWI_OK;
}

// Times `reps` launches of every candidate launch shape (steps per launch,
// rows per lane, chunk width) of the forward and the reverse kernels on the
// random-boundary grid (and the sponge forward on its grid) with nslots slots
// and keeps the fastest per time step.  The launches run on zeroed fields,
// and everything is zeroed again afterwards, so no pass state changes.
int fwi_autotune(fwi_ctx *c, int32_t nslots, int32_t reps, double *fwd_us, double *rev_us)
{
    TRY(check_ready(c));
    if (nslots < 1 || nslots > c->cfg.max_batch) return fail(FWI_ERR_ARG, "bad slot count");
    if (reps < 1) reps = 1;
    const int T = T_of(c);
    const long nf = c->lay.tune_fill(nslots);
    for (int k = 0; k < 8; ++k) TRY(fill(c, c->d_f[k], nf));
    TRY(set_shot0(c, 0));
    const int vzs[3] = {1, 2, 4};
    double best[3] = {1e30, 1e30, 1e30};
    for (int kind = 0; kind < 3; ++kind) {
        // kind 2, the sponge forward, runs on its own grid with the receivers
        // sampled (step index past tlag); kind 0 at step T (no source); the
        // reverse kernel at step 0, receiver data injected
        if (kind == 2 && !c->grid[0].on) continue;      // kernels of a grid that is not
        if (kind != 2 && !c->grid[1].on) continue;      // configured are not tuned
        const GridBuf &gb = c->grid[kind == 2 ? 0 : 1];
        const int tstep = (kind == 2 && c->tlag + 3 <= T) ? c->tlag + 1 : T;
        const long cells = (long)nslots * gb.g.nzz * gb.g.nxx;
        const int ncols = gb.g.nxx - 8;
        // times `n` launches of one shape on the batch; microseconds per step
        auto time_shape = [&](int st, int vz, int xwv, int n, double *us) -> int {
            c->nsteps[kind] = st;
            c->vz[kind] = vz;
            c->xw[kind] = xwv;
            fwi::Tiling tl = tiling_of(c, gb.g, kind);
            tl.staged = staging_fits(gb, tl, st, kind != 0, kind == 1);
            fwi::FwdStep fa = kind == 2 ? fwd_args(c, FWI_BOUNDARY_ABC, true)
                                        : fwd_args(c, FWI_BOUNDARY_RBC, false);
            fa.cur = c->d_f[1]; fa.old = c->d_f[0];
            fa.out1 = (st >= 2) ? c->d_f[2] : c->d_f[0]; fa.out2 = c->d_f[3];
            // receiver data injected at every step, as in a pass: the
            // slot-indexed seismogram buffer covers every slot
            fwi::RevStep ra = rev_args(c, 0);
            ra.data = c->d_seis;
            ra.data_by_shot = 0;
            ra.scur = c->d_f[1]; ra.sold = c->d_f[0];
            ra.sout1 = (st >= 2) ? c->d_f[2] : c->d_f[0]; ra.sout2 = c->d_f[3];
            ra.rcur = c->d_f[5]; ra.rold = c->d_f[4];
            ra.rout1 = (st >= 2) ? c->d_f[6] : c->d_f[4]; ra.rout2 = c->d_f[7];
            auto launch = [&]() {
                return kind != 1 ? fwi::launch_fwd_step(fa, tl, nslots, tstep, st, c->stream)
                                 : fwi::launch_rev_step(ra, tl, nslots, 0, st, c->stream);
            };
            for (int w = 0; w < 2; ++w) HIP_TRY(launch());
            HIP_TRY(hipEventRecord(c->ev0, c->stream));
            for (int r = 0; r < n; ++r) HIP_TRY(launch());
            HIP_TRY(hipEventRecord(c->ev1, c->stream));
            HIP_TRY(hipEventSynchronize(c->ev1));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            *us = 1e3 * ms / n / st;
            return FWI_OK;
        };
        struct Cand {
            double us;
            int st, vz, xw;
        };
        std::vector<Cand> cands;
        for (int st = 1; st <= 4; ++st) {
            // the forward kernels on LDS tiles (one shape per step count)
            for (int cpw : {16, 8, 4}) {            // LDS tiles: tile columns per wave
                double us = 0.0;
                TRY(time_shape(st, fwi::kTileVz, cpw, reps, &us));
                cands.push_back({us, st, fwi::kTileVz, cpw});
            }
            for (int a = 0; a < 3; ++a) {
                if (kind == 1 && st >= 3 && vzs[a] == 4) continue;
                // fixed widths, narrow ones for small batches (more waves in
                // flight), and widths that make the launch about 1-4 full
                // rounds of the waves the device holds at once
                std::vector<int> xws = {16, 32, 64, 96, 128};
                if (cells < (16L << 20)) { xws.push_back(4); xws.push_back(8); }
                const int tiles_per_chunk = fwi::zsegments(gb.g.nzz, vzs[a], st) * nslots;
                const int resident = fwi::resident_waves(kind, st, vzs[a]);
                for (int k = 1; k <= 4 && resident > 0; ++k) {
                    const int nchunk = (k * resident) / tiles_per_chunk;
                    if (nchunk < 1) continue;
                    const int xw = (ncols + nchunk - 1) / nchunk;
                    if (xw >= 4 && xw <= 256 &&
                        std::find(xws.begin(), xws.end(), xw) == xws.end())
                        xws.push_back(xw);
                }
                for (int xwv : xws) {
                    double us = 0.0;
                    TRY(time_shape(st, vzs[a], xwv, reps, &us));
                    cands.push_back({us, st, vzs[a], xwv});
                }
            }
        }
        // the three fastest again, with three times the launches: shapes
        // within a few per cent of each other are told apart by the longer run
        std::sort(cands.begin(), cands.end(),
                  [](const Cand &x, const Cand &y) { return x.us < y.us; });
        int bst = c->nsteps[kind], bvz = c->vz[kind], bxw = c->xw[kind];
        for (size_t n = 0; n < cands.size() && n < 3; ++n) {
            double us = 0.0;
            TRY(time_shape(cands[n].st, cands[n].vz, cands[n].xw, 3 * reps, &us));
            if (us < best[kind]) {
                best[kind] = us;
                bst = cands[n].st;
                bvz = cands[n].vz;
                bxw = cands[n].xw;
            }
        }
        c->nsteps[kind] = bst;
        c->vz[kind] = bvz;
        c->xw[kind] = bxw;
    }
    drop_graphs(c);
    for (int k = 0; k < 8; ++k) TRY(fill(c, c->d_f[k], nf));
    if (c->grid[1].on) {
        // the reverse launches accumulated into the per-slot image and
        // illumination of the random-boundary grid
        TRY(fill(c, c->d_acc_slot, c->lay.acc_fill(nslots)));
        TRY(fill(c, c->d_ill_slot, c->lay.acc_fill(nslots)));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (fwd_us) *fwd_us = c->grid[1].on ? best[0] : 0.0;
    if (rev_us) *rev_us = c->grid[1].on ? best[1] : 0.0;
    return FWI_OK;
}

// Device copy bandwidth of a plain float4 copy kernel over nbytes.
int fwi_stream_copy(int64_t nbytes, int32_t reps, double *gbs)
{
    if (nbytes < 1024 || reps < 1 || !gbs) return fail(FWI_ERR_ARG, "bad arguments");
    float *a = nullptr, *b = nullptr;
    const long n = (long)(nbytes / 16) * 4;
    TRY(dalloc(&a, n, nullptr));
    int r = dalloc(&b, n, nullptr);
    if (r != FWI_OK) { (void)hipFree(a); return r; }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipError_t e = fwi::launch_copy(a, b, n, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    for (int k = 0; k < reps && e == hipSuccess; ++k) e = fwi::launch_copy(a, b, n, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(FWI_ERR_HIP, "stream copy: %s", hipGetErrorString(e));
    *gbs = 2.0 * n * 4.0 * reps / (ms * 1e-3) / 1e9;
    return FWI_OK;
}

int fwi_profile_enable(fwi_ctx *c, int32_t enable)
{
    if (!c) return fail(FWI_ERR_ARG, "null context");
    c->profile = enable != 0;
    for (int k = 0; k < 2; ++k) {
        c->prof_ms[k] = 0.0;
        c->prof_launches[k] = 0;
        c->prof_cells[k] = 0.0;
    }
    return FWI_OK;
}

int fwi_profile_read(fwi_ctx *c, int32_t kind, double *ms, int64_t *launches, double *cells)
{
    if (!c || kind < 0 || kind > 1) return fail(FWI_ERR_ARG, "bad argument");
    if (ms) *ms = c->prof_ms[kind];
    if (launches) *launches = c->prof_launches[kind];
    if (cells) *cells = c->prof_cells[kind];
    return FWI_OK;
}

}  // extern "C"
