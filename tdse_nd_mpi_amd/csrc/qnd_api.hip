// qnd_api.hip -- libqnd: C ABI over the gfx950 split-operator passes.
//
// Internal representation.  For dim >= 2 a state between passes is held as
// M = FFT_0(psi): axis 0 in Fourier space, the other axes in position space.
// Then one split-operator stage is exactly two HBM round trips:
//   axis-0 pass : IFFT_0 -> (all axes in r) potential/CAP/field/mask/obs -> FFT_0
//   kinetic block: FFT over axes 1..d-1 -> (all axes in k) kinetic -> inverse
// and the kinetic factors of consecutive stages share one block.  For dim = 1
// the axis-0 pass also carries the kinetic factors (KPRE/KPOST), so a stage
// is one pass.  Slots are converted to position space lazily (download,
// observables that need full transforms).
#include "qnd_device.hpp"
#include "../../include/qnd.h"

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace qnd;

#define QND_VERSION "libqnd 0.1 (gfx950)"

namespace {

enum Rep { REP_R = 0, REP_M = 1 };

struct SlotState {
    int rep = REP_R;
    bool pend_mask = false;  // mask owed after the last real-time step
    bool pend_vend = false;  // velocity gauge: last V half-step owed
};

struct AxisPlan {
    int n = 0, log2n = 0, T = 0;
    void *tw = nullptr;  // device twiddle table, C<R>[n]
};

enum ProfClass { PC_POT = 0, PC_KIN = 1, PC_PLAIN = 2, PC_SMALL = 3, PC_BND = 4, PC_N = 5 };

}  // namespace

struct qnd_ctx {
    int dim = 0;
    int64_t nr[QND_MAXD] = {1, 1, 1, 1};
    double dr[QND_MAXD] = {0, 0, 0, 0};
    int64_t N = 0;
    int prec = QND_C128;
    size_t esz = 16;                      // bytes per stored complex value
    size_t lds_esz = 16;                  // bytes per complex value in the passes' registers and LDS
    bool x64 = false;                     // QND_C64F64: complex64 storage, FP64 pass arithmetic
    int n_slots = 0;
    int device = 0;
    size_t lds_max = 65536;
    int n_cu = 256;              // compute units of the device
    std::map<uint64_t, int> occ; // resident workgroups per CU of a pass shape (occupancy calculator)
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // second stream (default priority): adjoint recompute overlap
    Phys P{};
    bool have_pot = false, have_int = false;
    int s = 0;
    double as[20] = {0}, as_scaled[20] = {0};
    double dt = 0;
    double mask_frac = 0;
    void *d_states = nullptr;
    void *d_scr[3] = {nullptr, nullptr, nullptr};
    double *d_part = nullptr;
    size_t part_cap = 0;
    double *d_small = nullptr;  // coef [0,4096), gram [4096, 4096+8192), sums [12288, 16384)
    double *d_rec = nullptr;
    size_t rec_cap = 0;
    double *d_prmax = nullptr;
    double *d_capt = nullptr;    // per-axis CAP table (sum over axes = Vcap)
    double *d_maskt = nullptr;   // per-axis mask table (product over axes)
    int tw_override[2] = {0, 0}; // lines per workgroup (contiguous, strided); QND_TW env
    int vpt_override = 0;        // values per thread (4 or 8); QND_VPT env
    int wave_lines = 1;          // wavefront-local exchanges for 64-thread lines; QND_WAVE_LINES=0 disables
    bool round32 = false;        // QND_ROUND_STATE=1: complex128 passes store complex64-rounded states
    std::vector<SlotState> slots;
    AxisPlan plan[QND_MAXD];
    std::vector<double> Et, At, Eobs, Aobs;
    int64_t ntime = 0;
    int field_s = 0;             // stage count the field tables were laid out for
    double *d_bEt = nullptr, *d_bAt = nullptr, *d_bEo = nullptr;  // batch field tables (device)
    int bfield_s = 0;            // stage count of the batch tables
    // adjoint: checkpoints, saved position-space states, partials
    void *d_adj_ckpt = nullptr, *d_adj_final = nullptr, *d_adj_save = nullptr;
    double *d_adj_part = nullptr;
    size_t adj_ckpt_cap = 0, adj_save_cap = 0, adj_part_cap = 0;
    std::vector<SlotState> adj_ckpt_state;
    std::vector<double> adj_obs;
    int adj_slot = -1, adj_prec = -1;
    int64_t setup_gen = 0, adj_gen = -1;  // potential/integrator/mask/field changes; the forward's value
    int64_t adj_nsteps = 0, adj_K = 0;
    double *h_gs = nullptr;               // pinned: Gram-Schmidt flag, norm and step, two chunks
    hipEvent_t ev_gs[2] = {nullptr, nullptr};
    // tabulated kinetic factors (length gauge, four values per thread): one
    // table of N entries per distinct kinetic coefficient of a step
    void *d_ktab = nullptr;
    double *d_kcoef = nullptr;
    size_t ktab_cap = 0;
    std::vector<double> ktab_coef;
    int64_t ktab_gen = -1;
    int ktab_prec = -1;
    // per-axis kinetic factors of the step's first coefficient (merged step
    // boundary), E(t_n) on the device, and the ring of saved record states
    void *d_atab = nullptr;
    int64_t atab_gen = -1;
    // FP64 potential at every element (four-value passes), rebuilt when the
    // potential changes
    double *d_vtab = nullptr;
    int64_t pot_gen = 0, vtab_gen = -1;
    // per-axis kinetic factors of every kinetic coefficient of a step (phase:
    // s + 1 coefficients; imaginary time: s), Sigma n_j entries each
    void *d_kax = nullptr;
    std::vector<std::pair<int, double>> kax_coef;
    int64_t kax_gen = -1;
    int kax_prec = -1;
    bool use_kax = true;                  // QND_KAX=0: kinetic factors formed per element instead
    bool imag1d = true;                   // QND_IMAG1D=0: 1-D imaginary time in one launch per pass instead
    int gs_split = -1;                    // QND_GS_SPLIT=0/1: Gram-Schmidt fused / split (-1: by size)
    // descriptors of the group launches this context leads (qnd_group_real_steps)
    void *d_gdesc = nullptr;
    size_t gdesc_cap = 0;
    int atab_prec = -1;
    double *d_Eo = nullptr;
    void *d_side = nullptr;
    size_t side_cap = 0;  // bytes
    int bnb = 0;
    int64_t bntime = 0;
    hipEvent_t ev[16] = {};
    bool prof = false;
    int64_t prof_count[PC_N] = {0};
    double prof_ms[PC_N] = {0};
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_pending;
    std::string err;
};

static std::string g_create_err;

// ---------------------------------------------------------------------
// Scratch of the allocation-free entry point (qnd_observables_noalloc): the
// per-workgroup FP64 partials of its reduction and their column sums, in a
// file-scope device array that each GPU holds until the process ends.  In
// host code `noalloc_scratch` names a host-side stand-in, not the device
// array: the device address comes from hipGetSymbolAddress, once per device
// (qnd_create, with that device current), and only that address is used.
// ---------------------------------------------------------------------
#define QND_NOALLOC_ROWS 8192                                   // partial rows (workgroups) at most
#define QND_NOALLOC_DOUBLES (QND_NOALLOC_ROWS * QND_NOBS + 64)  // + the column sums
static __device__ __align__(256) double noalloc_scratch[QND_NOALLOC_DOUBLES];
#define QND_MAX_DEVICES 64
static std::mutex g_noalloc_mu;
static double *g_noalloc_ptr[QND_MAX_DEVICES] = {};
static hipError_t g_noalloc_err[QND_MAX_DEVICES] = {};
static bool g_noalloc_done[QND_MAX_DEVICES] = {};

// The device address of this device's scratch, looked up on the device's
// first context (the caller has made the device current).
static void noalloc_setup(int device)
{
    if (device < 0 || device >= QND_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_done[device]) return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc_scratch));
    g_noalloc_err[device] = e;
    g_noalloc_ptr[device] = e == hipSuccess ? (double *)ptr : nullptr;
    g_noalloc_done[device] = true;
}

namespace {

const int SMALL_COEF = 0, SMALL_GRAM = 4096, SMALL_SUMS = 12288, SMALL_STEP = 16000, SMALL_TOTAL = 16384;

int fail(qnd_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else g_create_err = buf;
    return code;
}

#define HIPCK(c, call)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail((c), QND_EHIP, "HIP error %s at %s:%d (%s)", hipGetErrorString(e_),  \
                        __FILE__, __LINE__, #call);                                           \
    } while (0)

// HIPCK for functions that return a pointer: sets *rcp and returns nullptr
#define HIPCK_P(c, call, rcp)                                                                 \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            *(rcp) = fail((c), QND_EHIP, "HIP error %s at %s:%d (%s)", hipGetErrorString(e_), \
                          __FILE__, __LINE__, #call);                                         \
            return nullptr;                                                                   \
        }                                                                                     \
    } while (0)

// Workgroups (256 threads) of the grid-stride flat kernels over N points:
// about `per` points per thread, clamped to [lo, hi], so that large grids
// fill every SIMD with several wavefronts (a fixed 512-1024 groups left an
// N = 512^3 reduction at 1 TB/s).  A reduction's order, hence its rounding,
// is fixed for a given N.
int flat_grid(int64_t N, int per, int lo, int hi)
{
    const int64_t g = N / (256 * (int64_t)per);
    return (int)(g < lo ? lo : g > hi ? hi : g);
}

int ilog2(int64_t n)
{
    int b = 0;
    while (((int64_t)1 << b) < n) ++b;
    return b;
}

template <typename R> C<R> *state_ptr(qnd_ctx *c, int slot)
{
    return reinterpret_cast<C<R> *>(c->d_states) + (int64_t)slot * c->N;
}

int ensure_part(qnd_ctx *c, size_t ndoubles)
{
    if (ndoubles <= c->part_cap) return 0;
    if (c->d_part) HIPCK(c, hipFree(c->d_part));
    size_t cap = ndoubles < 4096 ? 4096 : ndoubles;
    HIPCK(c, hipMalloc(&c->d_part, cap * sizeof(double)));
    c->part_cap = cap;
    return 0;
}

// ---- profiling wrapper ----
struct ProfScope {
    qnd_ctx *c;
    int cls;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(qnd_ctx *c_, int cls_) : c(c_), cls(cls_)
    {
        if (c->prof) {
            hipEventCreate(&a);
            hipEventCreate(&b);
            hipEventRecord(a, c->stream);
        }
    }
    ~ProfScope()
    {
        if (c->prof) {
            hipEventRecord(b, c->stream);
            c->prof_pending.push_back({cls, {a, b}});
        }
    }
};

void prof_flush(qnd_ctx *c)
{
    if (c->prof_pending.empty()) return;
    hipStreamSynchronize(c->stream);
    for (auto &p : c->prof_pending) {
        float ms = 0;
        hipEventElapsedTime(&ms, p.second.first, p.second.second);
        c->prof_count[p.first] += 1;
        c->prof_ms[p.first] += ms;
        hipEventDestroy(p.second.first);
        hipEventDestroy(p.second.second);
    }
    c->prof_pending.clear();
}

// ---- pass geometry ----
struct Geo {
    int axis, n, log2n, vpt, T, log2T, TW, log2TW, threads, grid, ldsStride;
    int64_t S, nlines;
    size_t lds;
};

int pow2floor(int64_t x)
{
    int p = 1;
    while ((int64_t)p * 2 <= x) p *= 2;
    return p;
}

// Workgroups of a strided-line pass kernel resident per CU with `threads`
// threads and `lds` bytes of LDS, from HIP's occupancy calculator (VGPRs,
// LDS and wavefront slots), cached per shape.  Axis 0 of a grid of two or
// more dimensions is the potential pass; other strided axes are plain passes.
template <typename R, int KIND, int DIR> const void *pass_fn(int vpt, bool x64 = false)
{
    if constexpr (std::is_same<R, float>::value) {
        if (x64)
            return vpt == 4 ? (const void *)k_pass_x<KIND, DIR, 4>
                   : vpt == 12 ? (const void *)k_pass_x<KIND, DIR, 12> : (const void *)k_pass_x<KIND, DIR, 8>;
    }
    return vpt == 4 ? (const void *)k_pass<R, KIND, DIR, 4>
           : vpt == 12 ? (const void *)k_pass<R, KIND, DIR, 12> : (const void *)k_pass<R, KIND, DIR, 8>;
}

// Workgroups per CU of a pass kernel of kind `kind` (PK_PLAIN, PK_KIN or PK_POT).
template <typename R> int blocks_per_cu(qnd_ctx *c, int kind, int vpt, int threads, size_t lds)
{
    const uint64_t key = ((uint64_t)lds << 24) | ((uint64_t)threads << 8) | ((uint64_t)vpt << 2) | (uint64_t)kind;
    auto it = c->occ.find(key);
    if (it != c->occ.end()) return it->second;
    const void *f = kind == PK_POT ? pass_fn<R, PK_POT, 1>(vpt, c->x64)
                    : kind == PK_KIN ? pass_fn<R, PK_KIN, -1>(vpt, c->x64) : pass_fn<R, PK_PLAIN, 1>(vpt, c->x64);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, threads, lds) != hipSuccess) {
        (void)hipGetLastError();
        nb = 0;
    }
    c->occ[key] = nb;
    return nb;
}

// Launch geometry of a pass along `axis` over nb consecutive slots (a batch
// of instances: an extra, untransformed leading axis).  `kind` is the pass
// kernel whose occupancy sizes strided tiles (-1: the potential pass on axis
// 0 of a grid of dim >= 2, else the plain pass).  `tuned` false ignores the
// tuning overrides, which are chosen for the passes of the run's own slots
// (the batched observable pass over saved record states takes the
// automatic geometry).
Geo geometry(qnd_ctx *c, int axis, int nb = 1, int kind = -1, bool tuned = true)
{
    const int vpt_ov = tuned ? c->vpt_override : 0;
    Geo g{};
    g.axis = axis;
    g.n = (int)c->nr[axis];
    g.S = 1;
    for (int j = axis + 1; j < c->dim; ++j) g.S *= c->nr[j];
    const int64_t lpi = c->N / g.n;  // lines per instance
    g.nlines = lpi * nb;
    if (g.n % 3 == 0) {
        // n = 3 * 2^k: twelve values per thread (radix 4/2 stages, radix 3 last)
        g.vpt = 12;
        g.log2n = ilog2(g.n / 3);
    } else {
        // values per thread: 8 (radix-8 stages) unless the grid is too small
        // to give every SIMD a wavefront that way, then 4 (twice the wavefronts)
        g.log2n = ilog2(g.n);
        g.vpt = 8;
        const bool few_waves = g.nlines * g.n / 8 / 64 < 4096;
        if ((vpt_ov == 4 || (vpt_ov == 0 && few_waves)) && g.n / 4 <= QND_PASS_THREADS)
            g.vpt = 4;
    }
    g.T = g.n / g.vpt;
    const size_t line_bytes = 2 * (size_t)g.n * c->lds_esz;  // two exchange buffers per line
    int twmax_lds = (int)(c->lds_max / line_bytes);
    if (twmax_lds < 1) twmax_lds = 1;
    int tw;
    if (g.S == 1) {
        tw = pow2floor(g.nlines / 1024 > 0 ? g.nlines / 1024 : 1);
        int cap = 256 / g.T > 0 ? 256 / g.T : 1;
        if (tw > cap) tw = cap;
        if (tw * g.T < 64) tw = 64 / g.T;  // at least one full wavefront
    } else {
        const int twmin = c->esz == 8 ? 4 : 2, twmaxseg = c->esz == 8 ? 16 : 8;
        tw = pow2floor(g.nlines / 256 > 0 ? g.nlines / 256 : 1);
        if (tw < twmin) tw = twmin;
        if (tw > twmaxseg) tw = twmaxseg;
        while (tw > 1 && (g.S % tw) != 0) tw /= 2;
    }
    const int ov = tuned ? c->tw_override[g.S == 1 ? 0 : 1] : 0;
    if (ov > 0) tw = pow2floor(ov);
    int thrcap = QND_PASS_THREADS / g.T;
    if (tw > thrcap) tw = thrcap;
    if (tw > twmax_lds) tw = pow2floor(twmax_lds);
    if ((int64_t)tw > g.nlines) tw = (int)g.nlines;
    if (g.S > 1 && (int64_t)tw > g.S) tw = (int)g.S;
    if (tw < 1) tw = 1;
    while (tw > 1 && lpi % tw != 0) tw /= 2;  // a workgroup never straddles two instances
    // one padding element per line spreads the strided tile rows over the banks
    while (tw > 1 && 2 * (size_t)tw * (g.n + 1) * c->lds_esz > c->lds_max) tw /= 2;
    if (g.S > 1 && ov == 0) {
        // Strided lines: of tw, tw/2 and tw/4 take the one that keeps the most
        // wavefronts resident on the chip (workgroups per CU from the kernel's
        // VGPRs and LDS, capped by the tiles there are), then the one that
        // uses more CUs, then the wider tile (longer row segments).  A kernel
        // with, e.g., 3 wavefronts per SIMD holds one 8-wavefront workgroup
        // per CU but three 4-wavefront ones.
        if (kind < 0) kind = (axis == 0 && c->dim > 1) ? PK_POT : PK_PLAIN;
        int64_t best_w = -1, best_cu = -1;
        int best = tw;
        for (int cand = tw; cand >= 1 && cand >= tw / 4; cand /= 2) {
            const int thr = g.T * cand;
            const size_t lds = 2 * (size_t)cand * (g.n + 1) * c->lds_esz;
            const int bpc = c->esz == 8 ? blocks_per_cu<float>(c, kind, g.vpt, thr, lds)
                                        : blocks_per_cu<double>(c, kind, g.vpt, thr, lds);
            if (bpc <= 0) continue;
            const int64_t tiles = g.nlines / cand;
            const int64_t resident = std::min<int64_t>(tiles, (int64_t)bpc * c->n_cu);
            const int64_t waves = resident * ((thr + 63) / 64);
            const int64_t cus = std::min<int64_t>(tiles, c->n_cu);
            if (waves > best_w || (waves == best_w && cus > best_cu)) {
                best_w = waves;
                best_cu = cus;
                best = cand;
            }
        }
        tw = best;
    }
    g.TW = tw;
    g.log2TW = ilog2(tw);
    g.log2T = ilog2(g.T);
    g.threads = g.T * tw;
    g.grid = (int)(g.nlines / tw);
    g.ldsStride = g.n + (tw > 1 ? 1 : 0);
    g.lds = 2 * (size_t)tw * g.ldsStride * c->lds_esz;
    return g;
}

template <typename R> PassArgs<R> base_args(qnd_ctx *c, C<R> *data, const Geo &g)
{
    PassArgs<R> a;
    memset(&a, 0, sizeof a);
    a.data = data;
    a.tw = reinterpret_cast<const typename TwT<R>::T *>(c->plan[g.axis].tw);
    a.axis = g.axis;
    a.n = g.n;
    a.log2n = g.log2n;
    a.T = g.T;
    a.log2T = g.log2T;
    a.TW = g.TW;
    a.log2TW = g.log2TW;
    a.ldsStride = g.ldsStride;
    a.wave_lines = c->wave_lines;
    a.kin_axes = (1u << c->dim) - 1u;
    a.S = g.S;
    a.nlines = g.nlines;
    a.scale1 = 1.0;
    a.scale2 = 1.0;
    a.P = c->P;
    a.slot_stride = c->N;
    a.B.nb = 1;
    a.B.lpi = c->N / g.n;
    return a;
}

// Launch of a pass.  When profiling, the kernel is launched with start/stop
// events recorded by the dispatch itself (hipExtLaunchKernelGGL), so the
// measured time is the kernel's own, as a kernel trace reports it.
template <typename R, int KIND, int DIR, int VPT>
void launch_vpt(qnd_ctx *c, const Geo &g, const PassArgs<R> &a, hipEvent_t e0, hipEvent_t e1)
{
    // leading scalars: what the addresses of the first loads need
    const unsigned geo = pass_geo(a.log2n, a.log2T, a.log2TW, a.P.pow2);
    // a finalizing kinetic pass appends one workgroup per instance after the
    // line tiles (whose order and XCD placement stay as they are)
    const int extra = (KIND == PK_KIN && (a.flags & PF_FINALIZE)) ? a.B.nb : 0;
    const dim3 grid(g.grid + extra);
    const unsigned fl = a.flags | (c->round32 ? (unsigned)PF_ROUND32 : 0u);
    if constexpr (std::is_same<R, float>::value) {
        if (c->x64) {
            if (e0)
                hipExtLaunchKernelGGL((k_pass_x<KIND, DIR, VPT>), grid, dim3(g.threads), g.lds, c->stream, e0, e1, 0,
                                      a.data, a.tw, a.S, geo, fl, (int)g.grid, a);
            else
                hipLaunchKernelGGL((k_pass_x<KIND, DIR, VPT>), grid, dim3(g.threads), g.lds, c->stream, a.data, a.tw,
                                   a.S, geo, fl, (int)g.grid, a);
            return;
        }
    }
    if (e0)
        hipExtLaunchKernelGGL((k_pass<R, KIND, DIR, VPT>), grid, dim3(g.threads), g.lds, c->stream, e0, e1, 0,
                              a.data, a.tw, a.S, geo, fl, (int)g.grid, a);
    else
        hipLaunchKernelGGL((k_pass<R, KIND, DIR, VPT>), grid, dim3(g.threads), g.lds, c->stream, a.data, a.tw,
                           a.S, geo, fl, (int)g.grid, a);
}

template <typename R, int KIND, int DIR>
int launch(qnd_ctx *c, const Geo &g, const PassArgs<R> &a, int cls)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->prof) {
        HIPCK(c, hipEventCreate(&e0));
        HIPCK(c, hipEventCreate(&e1));
        c->prof_pending.push_back({cls, {e0, e1}});
    }
    if (g.vpt == 4) launch_vpt<R, KIND, DIR, 4>(c, g, a, e0, e1);
    else if (g.vpt == 12) launch_vpt<R, KIND, DIR, 12>(c, g, a, e0, e1);
    else launch_vpt<R, KIND, DIR, 8>(c, g, a, e0, e1);
    HIPCK(c, hipGetLastError());
    return 0;
}

// The FP64 potential table of the latency-bound grids (four values per
// thread): read with the state at kernel entry instead of evaluating sqrt and
// exp in the pass.  Not built above QND_VTAB_MAX_BYTES (those grids run eight
// values per thread anyway).
#define QND_VTAB_MAX_BYTES ((size_t)256 << 20)
const double *vtab_for(qnd_ctx *c, int *rc)
{
    *rc = 0;
    if ((size_t)c->N * sizeof(double) > QND_VTAB_MAX_BYTES || !c->have_pot) return nullptr;
    if (c->vtab_gen == c->pot_gen && c->d_vtab) return c->d_vtab;
    if (!c->d_vtab) HIPCK_P(c, hipMalloc(&c->d_vtab, (size_t)c->N * sizeof(double)), rc);
    {
        ProfScope ps(c, PC_SMALL);
        hipLaunchKernelGGL(k_vtab, dim3(flat_grid(c->N, 4, 64, 4096)), dim3(256), 0, c->stream, c->d_vtab, c->N,
                           c->P);
        HIPCK_P(c, hipGetLastError(), rc);
    }
    c->vtab_gen = c->pot_gen;
    return c->d_vtab;
}

// axis-0 pass with r-space operators (IFFT first)
template <typename R> int launch_pot(qnd_ctx *c, const Geo &g, const PassArgs<R> &a_in)
{
    PassArgs<R> a = a_in;
    if (g.vpt == 4 && !a.vtab) {
        int rc;
        a.vtab = vtab_for(c, &rc);
        if (rc) return rc;
    }
    if (c->dim == 1) {
        if (a.flags == (PF_V1 | PF_V2 | PF_SECOND | PF_KPOST) && a.vmode == VM_IMAG && a.kmode == KM_IMAG)
            return launch<R, PK_P1DI, +1>(c, g, a, PC_POT);
        return launch<R, PK_POT1D, +1>(c, g, a, PC_POT);
    }
    if (a.flags == (PF_V1 | PF_SECOND) && a.vmode == VM_LENGTH) return launch<R, PK_POTL, +1>(c, g, a, PC_POT);
    if (a.flags == (PF_V1 | PF_V2 | PF_SECOND) && a.vmode == VM_IMAG) return launch<R, PK_POTI, +1>(c, g, a, PC_POT);
    if (a.flags == (PF_V1 | PF_V2 | PF_SECOND) && a.vmode == VM_VELOCITY)
        return launch<R, PK_POTV, +1>(c, g, a, PC_POT);
    if ((a.flags & ~PF_MASK) == (PF_V1 | PF_SECOND | PF_PREKX) && a.vmode == VM_LENGTH)
        return launch<R, PK_POTX, +1>(c, g, a, PC_BND);
    if (a.vmode == VM_IMAG && (a.flags & ~(PF_V1 | PF_V2 | PF_SECOND | PF_GS_UPDATE | PF_GS_DOTS)) == 0)
        return launch<R, PK_POTJ, +1>(c, g, a, PC_POT);
    if ((a.flags & ~PF_MASK) == (PF_OBS | PF_SECOND)) return launch<R, PK_POTB, +1>(c, g, a, PC_POT);
    return launch<R, PK_POT, +1>(c, g, a, PC_POT);
}

// plain transform of `data` along axis (dir -1 forward, +1 inverse), times scale
template <typename R> int plain_pass(qnd_ctx *c, C<R> *data, int axis, int dir, double scale, int nb = 1)
{
    Geo g = geometry(c, axis, nb);
    PassArgs<R> a = base_args<R>(c, data, g);
    a.scale1 = scale;
    if (dir < 0) return launch<R, PK_PLAIN, -1>(c, g, a, PC_PLAIN);
    return launch<R, PK_PLAIN, +1>(c, g, a, PC_PLAIN);
}

// full N-D transform (unnormalised)
template <typename R> int full_fft(qnd_ctx *c, C<R> *data, int dir)
{
    for (int ax = 0; ax < c->dim; ++ax) {
        int rc = plain_pass<R>(c, data, ax, dir, 1.0);
        if (rc) return rc;
    }
    return 0;
}

// Kinetic factor tables.  A length-gauge step (and its adjoint) uses s + 1
// kinetic coefficients: a_0, a_{2i+1} + a_{2i+2} (i < s-1) and a_{2s-1}.  For
// grids where a pass has four values per thread (one wavefront per SIMD,
// latency-bound) the factors e^{-i c K} of every element are tabulated once
// per set-up, so the kinetic pass reads its factor with the state instead
// of forming it.  Tables above QND_KTAB_MAX_BYTES are not built (large
// grids are not latency-bound, and the extra read would cost bandwidth).
#define QND_KTAB_MAX_BYTES ((size_t)256 << 20)
// ring of saved record states of the merged step boundary
#define QND_SIDE_MAX_BYTES ((size_t)256 << 20)
template <typename R>
const typename TwT<R>::T *ktab_for(qnd_ctx *c, double ck, const Geo &g, int *rc)
{
    using T = typename TwT<R>::T;
    *rc = 0;
    // the table holds the all-axes factor: 2-D only, where the kinetic block
    // is one pass
    if (g.vpt != 4 || c->dim != 2 || !c->have_int) return nullptr;
    if (c->ktab_gen != c->setup_gen || c->ktab_prec != c->prec) {
        c->ktab_gen = c->setup_gen;
        c->ktab_prec = c->prec;
        c->ktab_coef.clear();
        const int s = c->s;
        const double *as = c->as_scaled;
        std::vector<double> coef{as[0]};
        for (int i = 0; i < s - 1; ++i) coef.push_back(as[2 * i + 1] + as[2 * i + 2]);
        coef.push_back(as[2 * s - 1] + 0.0);
        const size_t bytes = coef.size() * (size_t)c->N * sizeof(T);
        if (bytes > QND_KTAB_MAX_BYTES) return nullptr;
        if (bytes > c->ktab_cap) {
            if (c->d_ktab) HIPCK_P(c, hipFree(c->d_ktab), rc);
            c->d_ktab = nullptr;
            c->ktab_cap = 0;
            HIPCK_P(c, hipMalloc(&c->d_ktab, bytes), rc);
            c->ktab_cap = bytes;
        }
        if (!c->d_kcoef) HIPCK_P(c, hipMalloc(&c->d_kcoef, 64 * sizeof(double)), rc);
        HIPCK_P(c, hipMemcpyAsync(c->d_kcoef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice,
                                  c->stream), rc);
        {
            ProfScope ps(c, PC_SMALL);
            hipLaunchKernelGGL(k_ktab<T>, dim3(flat_grid(c->N, 4, 64, 4096)), dim3(256), 0, c->stream,
                               (T *)c->d_ktab, (const double *)c->d_kcoef, (int)coef.size(), c->N, c->P);
            HIPCK_P(c, hipGetLastError(), rc);
        }
        c->ktab_coef = coef;
    }
    for (size_t q = 0; q < c->ktab_coef.size(); ++q)
        if (c->ktab_coef[q] == ck) return (const T *)c->d_ktab + q * c->N;
    return nullptr;
}

// Per-axis kinetic factors e^{-i a_0 p_j(i)^2 / 2} (a_0 the step's first
// kinetic coefficient, scaled by dt), Sigma n_j entries at the per-axis table
// offsets, formed in FP64 on the host (complex64: FP32 hi + lo).  Rebuilt
// when the set-up or the precision changes.
template <typename R> const typename TwT<R>::T *atab_for(qnd_ctx *c, int *rc)
{
    using T = typename TwT<R>::T;
    *rc = 0;
    if (c->atab_gen == c->setup_gen && c->atab_prec == c->prec && c->d_atab) return (const T *)c->d_atab;
    const int tot = c->P.off[c->dim - 1] + c->P.n[c->dim - 1];
    std::vector<T> h(tot);
    const double a0 = c->as_scaled[0];
    for (int j = 0; j < c->dim; ++j)
        for (int i = 0; i < c->P.n[j]; ++i) {
            const double p = (i < (c->P.n[j] >> 1)) ? i * c->P.dp[j] : (i - c->P.n[j]) * c->P.dp[j];
            const double ph = a0 * (.5 * p * p);
            const double cn = cos(ph), sn = -sin(ph);
            T &e = h[c->P.off[j] + i];
            if constexpr (std::is_same<T, double2>::value) {
                e.x = cn;
                e.y = sn;
            } else {
                e.x = (float)cn;
                e.y = (float)sn;
                e.z = (float)(cn - (double)e.x);
                e.w = (float)(sn - (double)e.y);
            }
        }
    if (c->d_atab) HIPCK_P(c, hipFree(c->d_atab), rc);
    c->d_atab = nullptr;
    HIPCK_P(c, hipMalloc(&c->d_atab, tot * sizeof(T)), rc);
    HIPCK_P(c, hipMemcpy(c->d_atab, h.data(), tot * sizeof(T), hipMemcpyHostToDevice), rc);
    c->atab_gen = c->setup_gen;
    c->atab_prec = c->prec;
    return (const T *)c->d_atab;
}

// Per-axis kinetic factors of every kinetic coefficient c of a step: the
// phase factors e^{-i c p_j(i)^2 / 2} of the real-time step's s + 1
// coefficients (a_0, a_{2i+1} + a_{2i+2}, a_{2s-1}) and the damping factors
// e^{-c p_j(i)^2 / 2} of the imaginary-time step's s coefficients
// (a_{2i} + a_{2i+1}), Sigma n_j entries per coefficient at the per-axis
// offsets.  Formed in FP64 on the host (complex64: FP32 hi + lo; a damping
// factor in the first component, its exponent clamped as in the pass).  The
// kinetic factor of a pass is then one table entry per element and one per
// folded axis of its line, instead of an angle reduction and sin/cos per
// element.  Rebuilt when the set-up or the precision changes; nullptr when c
// is not one of the step's coefficients (or QND_KAX=0).
template <typename R>
const typename TwT<R>::T *kax_for(qnd_ctx *c, int kmode, double ck, int *rc)
{
    using T = typename TwT<R>::T;
    *rc = 0;
    if (!c->use_kax || !c->have_int || (kmode != KM_PHASE && kmode != KM_IMAG)) return nullptr;
    const int tot = c->P.off[c->dim - 1] + c->P.n[c->dim - 1];
    if (c->kax_gen != c->setup_gen || c->kax_prec != c->prec || !c->d_kax) {
        const int s = c->s;
        const double *as = c->as_scaled;
        std::vector<std::pair<int, double>> coef{{KM_PHASE, as[0]}};
        for (int i = 0; i < s - 1; ++i) coef.push_back({KM_PHASE, as[2 * i + 1] + as[2 * i + 2]});
        coef.push_back({KM_PHASE, as[2 * s - 1] + 0.0});
        for (int i = 0; i < s; ++i) coef.push_back({KM_IMAG, as[2 * i] + as[2 * i + 1]});
        const double clamp = Prec<R>::clamp;
        std::vector<T> h(coef.size() * (size_t)tot);
        for (size_t q = 0; q < coef.size(); ++q)
            for (int j = 0; j < c->dim; ++j)
                for (int i = 0; i < c->P.n[j]; ++i) {
                    const double p = (i < (c->P.n[j] >> 1)) ? i * c->P.dp[j] : (i - c->P.n[j]) * c->P.dp[j];
                    const double K = .5 * p * p;
                    double cn, sn;
                    if (coef[q].first == KM_PHASE) {
                        const double ph = coef[q].second * K;
                        cn = cos(ph);
                        sn = -sin(ph);
                    } else {
                        const double arg = -coef[q].second * K;
                        cn = exp(arg > clamp ? clamp : arg);
                        sn = 0.0;
                    }
                    T &e = h[q * tot + c->P.off[j] + i];
                    if constexpr (std::is_same<T, double2>::value) {
                        e.x = cn;
                        e.y = sn;
                    } else {
                        e.x = (float)cn;
                        e.y = (float)sn;
                        e.z = (float)(cn - (double)e.x);
                        e.w = (float)(sn - (double)e.y);
                    }
                }
        if (c->d_kax) HIPCK_P(c, hipFree(c->d_kax), rc);
        c->d_kax = nullptr;
        HIPCK_P(c, hipMalloc(&c->d_kax, h.size() * sizeof(T)), rc);
        HIPCK_P(c, hipMemcpy(c->d_kax, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), rc);
        c->kax_gen = c->setup_gen;
        c->kax_prec = c->prec;
        c->kax_coef = coef;
    }
    for (size_t q = 0; q < c->kax_coef.size(); ++q)
        if (c->kax_coef[q].first == kmode && c->kax_coef[q].second == ck) return (const T *)c->d_kax + q * tot;
    return nullptr;
}

// Where the per-axis tables pay (measured): 2-D kinetic passes with eight or
// twelve values per thread (the 384^2 sweep class 3.5 % faster, 256^2
// unchanged).  Not with four values per thread, where one wavefront per SIMD
// would wait on the table read between the transforms, and not in 3-D, where
// the 512^3 kinetic passes became 7 % slower with 25 % fewer VALU
// instructions: they are bound by the state's memory latency, and the table
// read adds one more wait.
static bool kax_pays(const qnd_ctx *c, const Geo &g) { return c->dim == 2 && g.vpt != 4; }

// A pending observable record (partials of the previous axis-0 pass) is
// finalized by workgroups appended to the first kinetic pass when fin is given.
struct Finalize {
    const double *part;
    int nwg;         // workgroups per instance of the pass that wrote the partials
    double *rec;     // record of instance 0
    ObsSpec spec;
    Batch B;         // nb, per-instance E(t_n) and record stride
};

// Block size of the one-block column reductions over nwg partial rows: at
// least eight rows per thread, so small grids reduce within one wavefront
// and large ones use 1024 threads (the sums' order, hence their rounding, is
// fixed for a given nwg).
int reduce_threads(int64_t nwg)
{
    int thr = 64;
    while (thr < 1024 && (int64_t)thr * 8 < nwg) thr *= 2;
    return thr;
}

int launch_finalize(qnd_ctx *c, const Finalize *fin)
{
    ProfScope ps(c, PC_SMALL);
    hipLaunchKernelGGL(k_obs_finalize, dim3(fin->B.nb), dim3(reduce_threads(fin->nwg)), 0, c->stream, fin->part,
                       fin->nwg, c->dim,
                       c->P.dNr, fin->spec, fin->B, fin->rec);
    HIPCK(c, hipGetLastError());
    return 0;
}

// A per-instance field row for the passes of a batch: base of instance 0 and
// the stride between instances (nullptr: the scalar field of the call).
struct FieldRows {
    const double *base = nullptr;
    int64_t stride = 0;
};

// Kinetic block for dim >= 2, on the axis-0 Fourier representation M.  The
// kinetic factor exp(-i c sum_j K_j) (K_j = p_j^2 / 2, plus (p_j + A_j/2) A_j
// in the velocity gauge; exp(-c sum_j K_j) in imaginary time) is a product of
// commuting per-axis factors, so the block is one pass per axis 1 .. d-1:
// forward transform along the axis, the factor of its axes, inverse.  The
// pass along axis 1 also applies axis 0's factor (axis 0 is in k-space in M).
// In 2-D this is the single all-k pass; in 3-D two passes replace the forward
// transform, the all-k pass and the inverse transform (three passes).  Each
// factor's exponent is clamped separately (a clamp that would act signals a
// blown-up state in either form).
template <typename R>
int kinetic_block(qnd_ctx *c, C<R> *data, int kmode, double ck, const double *A, const Finalize *fin = nullptr,
                  int nb = 1, FieldRows Arows = FieldRows(), unsigned flags = 0)
{
    const int d = c->dim;
    int rc;
    for (int ax = 1; ax < d; ++ax) {
        Geo g = geometry(c, ax, nb, PK_KIN);
        PassArgs<R> a = base_args<R>(c, data, g);
        a.flags = flags;
        a.kmode = kmode;
        a.kin_axes = ax == 1 ? 3u : (1u << ax);
        a.ck_mid = ck;
        a.scale2 = 1.0 / (double)c->nr[ax];
        if (kmode == KM_PHASE) {
            a.ktab = ktab_for<R>(c, ck, g, &rc);
            if (rc) return rc;
            if (a.ktab) a.flags |= PF_KTAB;
        }
        if ((kmode == KM_PHASE || kmode == KM_IMAG) && !a.ktab && kax_pays(c, g)) {
            a.ktab = kax_for<R>(c, kmode, ck, &rc);
            if (rc) return rc;
            if (a.ktab) a.flags |= PF_KAX;
        }
        if (A)
            for (int j = 0; j < d; ++j) a.Ak[j] = A[j];
        a.B.nb = nb;
        a.B.fld = Arows.base;
        a.B.fld_stride = Arows.stride;
        if (fin && ax == 1) {
            a.flags |= PF_FINALIZE;
            a.fin_part = fin->part;
            a.fin_nwg = fin->nwg;
            a.fin_rec = fin->rec;
            a.fin_spec = fin->spec;
            a.B.finE = fin->B.finE;
            a.B.finE_stride = fin->B.finE_stride;
            a.B.rec_stride = fin->B.rec_stride;
        }
        if (kmode == KM_PHASE && !(a.flags & PF_FINALIZE)) {
            if ((rc = launch<R, PK_KINP, -1>(c, g, a, PC_KIN))) return rc;
        } else if (kmode == KM_IMAG && !(a.flags & PF_FINALIZE)) {
            if ((rc = launch<R, PK_KINI, -1>(c, g, a, PC_KIN))) return rc;
        } else if (kmode == KM_VELOCITY && !(a.flags & PF_FINALIZE)) {
            if ((rc = launch<R, PK_KINV, -1>(c, g, a, PC_KIN))) return rc;
        } else if ((rc = launch<R, PK_KIN, -1>(c, g, a, PC_KIN))) {
            return rc;
        }
    }
    return 0;
}

// Bring a slot to representation `want`, paying owed velocity/mask ops.
template <typename R> int ensure_rep(qnd_ctx *c, int slot, int want)
{
    SlotState &st = c->slots[slot];
    C<R> *psi = state_ptr<R>(c, slot);
    if (st.rep == want && !st.pend_mask && !st.pend_vend) return 0;
    if (st.rep == REP_M) {
        // M -> R with the owed operators
        Geo g = geometry(c, 0);
        PassArgs<R> a = base_args<R>(c, psi, g);
        a.scale1 = 1.0 / (double)c->nr[0];
        if (st.pend_vend) {
            a.flags |= PF_V1;
            a.vmode = VM_VELOCITY;
            a.cv1 = c->as_scaled[2 * c->s - 1];
        }
        if (st.pend_mask) a.flags |= PF_MASK;
        int rc = launch_pot<R>(c, g, a);
        if (rc) return rc;
        st.rep = REP_R;
        st.pend_mask = st.pend_vend = false;
    }
    if (want == REP_M && st.rep == REP_R) {
        int rc = plain_pass<R>(c, psi, 0, -1, 1.0);
        if (rc) return rc;
        st.rep = REP_M;
    }
    return 0;
}

// Pay every slot's owed operators (velocity-gauge last V half-step, mask)
// with the current set-up.  Called before the set-up changes, so that an
// operator owed by the last real-time steps is the one those steps used.
int flush_owed(qnd_ctx *c)
{
    for (int k = 0; k < c->n_slots; ++k) {
        const SlotState &st = c->slots[k];
        if (!st.pend_mask && !st.pend_vend) continue;
        const int rc = c->prec == QND_C64 ? ensure_rep<float>(c, k, REP_R) : ensure_rep<double>(c, k, REP_R);
        if (rc) return rc;
    }
    return 0;
}

// Reduction helper: launch k_red and sum partials into d_small[SMALL_SUMS..]
template <typename R>
int reduce(qnd_ctx *c, int mode, const C<R> *a, C<R> *b, const C<R> *cc, int axis, double fs,
           const double *alpha_dev, int nq, double scale, double *host_out)
{
    const int grid = flat_grid(c->N, 8, 512, 8192), npart = QND_NOBS;
    int rc = ensure_part(c, (size_t)grid * npart);
    if (rc) return rc;
    {
        ProfScope ps(c, PC_SMALL);
        hipLaunchKernelGGL(k_red<R>, dim3(grid), dim3(256), 0, c->stream, mode, a, b, cc, axis, fs,
                           c->N, c->P, alpha_dev, c->d_part, npart);
        HIPCK(c, hipGetLastError());
        hipLaunchKernelGGL(k_sum_partials, dim3((nq + 3) / 4), dim3(256), 0, c->stream, c->d_part, grid,
                           npart, nq, scale, c->d_small + SMALL_SUMS);
        HIPCK(c, hipGetLastError());
    }
    if (host_out) {
        HIPCK(c, hipMemcpyAsync(host_out, c->d_small + SMALL_SUMS, nq * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream));
        HIPCK(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}

template <typename R> int elem(qnd_ctx *c, int mode, C<R> *dst, const C<R> *src, int axis, double shift)
{
    ProfScope ps(c, PC_SMALL);
    hipLaunchKernelGGL(k_elem<R>, dim3(flat_grid(c->N, 4, 1024, 16384)), dim3(256),
                       0, c->stream, mode, dst, src, axis, shift,
                       c->N, c->P, (const C<R> *)nullptr, (int64_t)0, 0, (const double *)nullptr);
    HIPCK(c, hipGetLastError());
    return 0;
}

template <typename R> int copy_state(qnd_ctx *c, C<R> *dst, const C<R> *src)
{
    HIPCK(c, hipMemcpyAsync(dst, src, c->N * sizeof(C<R>), hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

// ---------------------------------------------------------------------
// energies (src/calculate_observables.c:146-175)
// ---------------------------------------------------------------------
template <typename R> int energies(qnd_ctx *c, int slot, int M, double *E)
{
    int rc = ensure_rep<R>(c, slot, REP_R);
    if (rc) return rc;
    C<R> *psi = state_ptr<R>(c, slot);
    C<R> *H = reinterpret_cast<C<R> *>(c->d_scr[0]);
    C<R> *A = reinterpret_cast<C<R> *>(c->d_scr[1]);
    double fs = 1.0;
    for (int j = 0; j < c->dim; ++j) fs /= (double)c->nr[j];
    if ((rc = copy_state<R>(c, H, psi))) return rc;
    for (int m = 0; m < M; ++m) {
        if ((rc = copy_state<R>(c, A, H))) return rc;
        if ((rc = full_fft<R>(c, A, -1))) return rc;
        if ((rc = elem<R>(c, EM_MUL_KIN, A, A, 0, 0.0))) return rc;
        if ((rc = full_fft<R>(c, A, +1))) return rc;
        double out[2];
        if ((rc = reduce<R>(c, RM_HCOMB, psi, H, A, 0, fs, nullptr, 2, c->P.dNr, out))) return rc;
        E[m] = out[0];
    }
    return 0;
}

// Lz (src/calculate_observables.c:178-230), dim >= 2
template <typename R> int lz(qnd_ctx *c, int slot, const double *alpha, double *out)
{
    int rc = ensure_rep<R>(c, slot, REP_R);
    if (rc) return rc;
    C<R> *psi = state_ptr<R>(c, slot);
    C<R> *Pp = reinterpret_cast<C<R> *>(c->d_scr[0]);
    C<R> *B = reinterpret_cast<C<R> *>(c->d_scr[1]);
    double fs = 1.0;
    for (int j = 0; j < c->dim; ++j) fs /= (double)c->nr[j];
    if ((rc = copy_state<R>(c, Pp, psi))) return rc;
    if ((rc = full_fft<R>(c, Pp, -1))) return rc;
    double r1[2], r2[2];
    if ((rc = elem<R>(c, EM_MUL_PSHIFT, B, Pp, 1, alpha ? alpha[1] : 0.0))) return rc;
    if ((rc = full_fft<R>(c, B, +1))) return rc;
    if ((rc = reduce<R>(c, RM_DOTX, psi, B, nullptr, 0, 0, nullptr, 2, 1.0, r1))) return rc;
    if ((rc = elem<R>(c, EM_MUL_PSHIFT, B, Pp, 0, alpha ? alpha[0] : 0.0))) return rc;
    if ((rc = full_fft<R>(c, B, +1))) return rc;
    if ((rc = reduce<R>(c, RM_DOTX, psi, B, nullptr, 1, 0, nullptr, 2, 1.0, r2))) return rc;
    *out = (r1[0] - r2[0]) * fs * c->P.dNr;
    return 0;
}

int obs_width(int code, int dim)
{
    switch (code) {
    case QND_OBS_LZ: case QND_OBS_ENERGY: case QND_OBS_NORM: return 1;
    case QND_OBS_DIPOLE_R: case QND_OBS_DIPOLE_V: case QND_OBS_DIPOLE_A: return dim;
    default: return 0;
    }
}

// Calculation_global_observables (src/calculate_observables.c:119-140)
template <typename R>
int observables(qnd_ctx *c, int slot, const int *codes, int ncodes, const double *alpha,
                const double *Efield, double *out)
{
    int rc = 0;
    int off = 0;
    const int d = c->dim;
    double zero[QND_MAXD] = {0, 0, 0, 0};
    if (!alpha) alpha = zero;
    if (!Efield) Efield = zero;
    for (int k = 0; k < ncodes; ++k) {
        const int code = codes[k];
        if (code == QND_OBS_LZ) {
            if (d <= 1) out[off] = 0.0;
            else if ((rc = lz<R>(c, slot, alpha, &out[off]))) return rc;
        } else if (code == QND_OBS_ENERGY) {
            if ((rc = energies<R>(c, slot, 1, &out[off]))) return rc;
        } else if (code == QND_OBS_DIPOLE_R || code == QND_OBS_NORM || code == QND_OBS_DIPOLE_A) {
            if ((rc = ensure_rep<R>(c, slot, REP_R))) return rc;
            C<R> *psi = state_ptr<R>(c, slot);
            double v[QND_NOBS];
            if ((rc = reduce<R>(c, RM_OBS, psi, nullptr, nullptr, 0, 0, nullptr, QND_NOBS, c->P.dNr, v)))
                return rc;
            if (code == QND_OBS_NORM) out[off] = v[0];
            else if (code == QND_OBS_DIPOLE_R)
                for (int j = 0; j < d; ++j) out[off + j] = v[1 + j];
            else
                for (int j = 0; j < d; ++j) out[off + j] = -v[1 + QND_MAXD + j] - Efield[j] * v[0];
        } else if (code == QND_OBS_DIPOLE_V) {
            if ((rc = ensure_rep<R>(c, slot, REP_R))) return rc;
            C<R> *psi = state_ptr<R>(c, slot);
            C<R> *A = reinterpret_cast<C<R> *>(c->d_scr[0]);
            if ((rc = copy_state<R>(c, A, psi))) return rc;
            if ((rc = full_fft<R>(c, A, -1))) return rc;
            double *dal = c->d_small + SMALL_SUMS + 64;
            HIPCK(c, hipMemcpyAsync(dal, alpha, d * sizeof(double), hipMemcpyHostToDevice, c->stream));
            double fs = 1.0, dNp = 1.0;
            for (int j = 0; j < d; ++j) {
                fs /= (double)c->nr[j];
                dNp *= 2. * M_PI / (double)(c->dr[j] * c->nr[j]);
            }
            double v[QND_MAXD];
            if ((rc = reduce<R>(c, RM_PMOM, A, nullptr, nullptr, 0, 0, dal, d, dNp * fs, v))) return rc;
            for (int j = 0; j < d; ++j) out[off + j] = v[j];
        }
        off += obs_width(code, d);
    }
    return 0;
}

// Gram matrix of slots 0..nprev-1 in the M representation, into d_small.
// G is Hermitian: the entries with i <= j are reduced, G_ji = conj(G_ij).
// Up to QND_GRAM_MAX states one kernel reads each state once and sums every
// pair; beyond that, one reduction per pair.
template <typename R, int NP> void launch_gram(qnd_ctx *c, int grid)
{
    hipLaunchKernelGGL((k_gram<R, NP>), dim3(grid), dim3(256), 0, c->stream, state_ptr<R>(c, 0), c->N, c->N,
                       c->d_part);
}
template <typename R> int gram(qnd_ctx *c, int nprev)
{
    std::vector<double> G(2 * (size_t)nprev * nprev, 0.0);
    const double scale = c->P.dNr / (double)c->nr[0];
    auto put = [&](int i, int j, const double *v) {
        G[2 * (i * nprev + j)] = v[0];
        G[2 * (i * nprev + j) + 1] = i == j ? 0.0 : v[1];
        G[2 * (j * nprev + i)] = v[0];
        G[2 * (j * nprev + i) + 1] = i == j ? 0.0 : -v[1];
    };
    if (nprev <= QND_GRAM_MAX) {
        const int npair = nprev * (nprev + 1) / 2, nq = 2 * npair;
        const int grid = flat_grid(c->N, 8, 512, 8192);
        int rc = ensure_part(c, (size_t)grid * nq);
        if (rc) return rc;
        std::vector<double> v(nq);
        {
            ProfScope ps(c, PC_SMALL);
            switch (nprev) {
            case 1: launch_gram<R, 1>(c, grid); break;
            case 2: launch_gram<R, 2>(c, grid); break;
            case 3: launch_gram<R, 3>(c, grid); break;
            case 4: launch_gram<R, 4>(c, grid); break;
            case 5: launch_gram<R, 5>(c, grid); break;
            case 6: launch_gram<R, 6>(c, grid); break;
            case 7: launch_gram<R, 7>(c, grid); break;
            default: launch_gram<R, 8>(c, grid); break;
            }
            HIPCK(c, hipGetLastError());
            hipLaunchKernelGGL(k_sum_partials, dim3((nq + 3) / 4), dim3(256), 0, c->stream, c->d_part, grid, nq, nq,
                               scale, c->d_small + SMALL_SUMS);
            HIPCK(c, hipGetLastError());
        }
        HIPCK(c, hipMemcpyAsync(v.data(), c->d_small + SMALL_SUMS, nq * sizeof(double), hipMemcpyDeviceToHost,
                                c->stream));
        HIPCK(c, hipStreamSynchronize(c->stream));
        int p = 0;
        for (int i = 0; i < nprev; ++i)
            for (int j = i; j < nprev; ++j, ++p) put(i, j, &v[2 * p]);
    } else {
        for (int i = 0; i < nprev; ++i)
            for (int j = i; j < nprev; ++j) {
                double v[2];
                int rc = reduce<R>(c, RM_DOT, state_ptr<R>(c, i), state_ptr<R>(c, j), nullptr, 0, 0, nullptr,
                                   2, scale, v);
                if (rc) return rc;
                put(i, j, v);
            }
    }
    HIPCK(c, hipMemcpyAsync(c->d_small + SMALL_GRAM, G.data(), G.size() * sizeof(double),
                            hipMemcpyHostToDevice, c->stream));
    return 0;
}

// ---------------------------------------------------------------------
// imaginary time (src/integrators.c:135-163 + src/QuantumND_imag.c:265-311)
// ---------------------------------------------------------------------
// Blocks of the first-level sums of the Gram-Schmidt dot partials: none for
// up to 4096 rows (the coefficient kernel sums them directly), else one per
// 128 rows, at most 256.
int gs_row_blocks(int nrows) { return nrows <= 4096 ? 0 : std::min(256, (nrows + 127) / 128); }

// Gram-Schmidt of large grids as two contiguous elementwise kernels
// (k_gs_dots after the step, k_gs_update before the next), launched for NP
// previous states (at most QND_GRAM_MAX) on GS_SPLIT_BLOCKS workgroups.
#define GS_SPLIT_BLOCKS 2048
#define QND_GS_SPLIT_MIN ((int64_t)1 << 24)  // grid points from which the split form is used
template <typename R, int NP> void launch_gs_split(qnd_ctx *c, bool dots, C<R> *psi, double *coef)
{
    const int grid = (int)std::min<int64_t>(GS_SPLIT_BLOCKS, (c->N + 255) / 256);
    if (dots)
        hipLaunchKernelGGL((k_gs_dots<R, NP>), dim3(grid), dim3(256), 0, c->stream, (const C<R> *)psi,
                           (const C<R> *)state_ptr<R>(c, 0), c->N, c->N, c->d_part);
    else
        hipLaunchKernelGGL((k_gs_update<R, NP>), dim3(grid), dim3(256), 0, c->stream, psi,
                           (const C<R> *)state_ptr<R>(c, 0), c->N, c->N, (const double *)coef);
}
template <typename R> void gs_split(qnd_ctx *c, int nprev, bool dots, C<R> *psi, double *coef)
{
    switch (nprev) {
    case 0: launch_gs_split<R, 0>(c, dots, psi, coef); break;
    case 1: launch_gs_split<R, 1>(c, dots, psi, coef); break;
    case 2: launch_gs_split<R, 2>(c, dots, psi, coef); break;
    case 3: launch_gs_split<R, 3>(c, dots, psi, coef); break;
    case 4: launch_gs_split<R, 4>(c, dots, psi, coef); break;
    case 5: launch_gs_split<R, 5>(c, dots, psi, coef); break;
    case 6: launch_gs_split<R, 6>(c, dots, psi, coef); break;
    case 7: launch_gs_split<R, 7>(c, dots, psi, coef); break;
    default: launch_gs_split<R, 8>(c, dots, psi, coef); break;
    }
}

// k_imag1d for a 1-D line of 2^log2n points in the automatic geometry's
// values per thread: 256 to 2048 points (four) and 4096 (eight); nullptr for
// the other lines (shorter, 3 * 2^k, or another geometry), which keep one
// launch per pass
template <typename R> const void *imag1d_fn(int vpt, int log2n)
{
#define QND_I1D(V, Lc)                                                                                              \
    if (vpt == V && log2n == Lc) return reinterpret_cast<const void *>(&k_imag1d<R, V, Lc>);
    QND_I1D(4, 8) QND_I1D(4, 9) QND_I1D(4, 10) QND_I1D(4, 11) QND_I1D(8, 12)
#undef QND_I1D
    return nullptr;
}

template <typename R> int imag_steps(qnd_ctx *c, int slot, int64_t nsteps, int gs)
{
    const int d = c->dim, s = c->s;
    const int nprev = gs ? slot : 0;
    int rc;
    for (int j = 0; j < nprev; ++j)
        if ((rc = ensure_rep<R>(c, j, REP_M))) return rc;
    if ((rc = ensure_rep<R>(c, slot, REP_M))) return rc;
    if (gs && nprev > 0 && (rc = gram<R>(c, nprev))) return rc;
    C<R> *psi = state_ptr<R>(c, slot);
    double *coef = c->d_small + SMALL_COEF;
    HIPCK(c, hipMemsetAsync(coef, 0, (2 * nprev + 4) * sizeof(double), c->stream));
    const Geo g0 = geometry(c, 0);
    const int npart = 2 * (nprev + 1);
    // Large grids (or QND_GS_SPLIT=1) take the split form: the dots and the
    // update as contiguous kernels of their own instead of fused into the
    // first and last pass of a step
    const bool split = gs && nprev <= QND_GRAM_MAX &&
                       (c->gs_split > 0 || (c->gs_split < 0 && c->N >= QND_GS_SPLIT_MIN));
    const int dot_rows = split ? (int)std::min<int64_t>(GS_SPLIT_BLOCKS, (c->N + 255) / 256) : g0.grid;
    // Gram-Schmidt dot partials: one row per workgroup that wrote them, and
    // (many rows) one row per block of their first-level sums behind them
    const int gs_blocks = gs_row_blocks(dot_rows);
    if (gs && (rc = ensure_part(c, (size_t)(dot_rows + gs_blocks) * npart))) return rc;
    const double inv_n0 = 1.0 / (double)c->nr[0];
    bool pending = false;
    // Gram-Schmidt collapse (src/QuantumND_imag.c:297-304): the coefficient
    // kernel records the step of the first collapse.  Every GS_CHUNK steps the
    // flag is copied to pinned host memory behind an event; the host reads the
    // previous chunk's copy after enqueuing the next chunk, so the check never
    // drains the queue, and a call stops at most two chunks after the step.
    const int64_t GS_CHUNK = 64;
    auto collapse_error = [&](const double *h) {
        return fail(c, QND_EGS_COLLAPSE,
                    "ERROR: Wavefunction norm decayed to zero during Gram-Schmidt for state %d "
                    "(step %lld of this call, norm %e, tolerance 1e-12).",
                    slot, (long long)h[2], h[1]);
    };
    int64_t chunks_out = 0;  // flag copies enqueued
    // 1-D: the whole call in one launch of one workgroup (k_imag1d), for the
    // line lengths it is compiled for
    const void *k1d = d == 1 && c->imag1d && !c->x64 && !c->prof && s <= 10 && nsteps > 0 &&
                              nprev <= QND_GRAM_MAX && g0.nlines == 1 && g0.TW == 1 && g0.T <= 512
                          ? imag1d_fn<R>(g0.vpt, g0.log2n)
                          : nullptr;
    const bool one_launch = k1d != nullptr;
    if (one_launch) {
        Imag1D A;
        memset(&A, 0, sizeof A);
        for (int k = 0; k < 2 * s; ++k) A.as[k] = c->as_scaled[k];
        A.s = s;
        A.nprev = gs ? nprev : -1;
        A.ldsStride = g0.ldsStride;
        A.inv_n = inv_n0;
        A.gs_scale = c->P.dNr * inv_n0;
        const auto *tw = reinterpret_cast<const typename TwT<R>::T *>(c->plan[0].tw);
        const size_t lds = 2 * (size_t)g0.ldsStride * sizeof(C<R>);
        const C<R> *prev = state_ptr<R>(c, 0);
        const double *gram = c->d_small + SMALL_GRAM;
        int64_t slot_stride = c->N;
        Phys P = c->P;
        void *args[] = {&psi, &prev, &slot_stride, &tw, &P, &A, &gram, &coef, &nsteps};
        HIPCK(c, hipLaunchKernel(k1d, dim3(1), dim3(g0.T), args, lds, c->stream));
    }
    for (int64_t it = 0; it < (one_launch ? 0 : nsteps); ++it) {
        for (int ia = 0; ia < s; ++ia) {
            PassArgs<R> a = base_args<R>(c, psi, g0);
            a.scale1 = inv_n0;
            a.flags = PF_V2 | PF_SECOND;
            a.vmode = VM_IMAG;
            a.cv2 = c->as_scaled[2 * ia];
            if (ia > 0) {
                a.flags |= PF_V1;
                a.cv1 = c->as_scaled[2 * ia - 1];
            }
            if (ia == 0 && pending) {
                if (split) {
                    ProfScope ps(c, PC_SMALL);
                    gs_split<R>(c, nprev, false, psi, coef);
                    HIPCK(c, hipGetLastError());
                } else {
                    a.flags |= PF_GS_UPDATE;
                    a.prev = state_ptr<R>(c, 0);
                    a.nprev = nprev;
                    a.gs_coef = coef;
                }
                pending = false;
            }
            const double ck = c->as_scaled[2 * ia] + c->as_scaled[2 * ia + 1];
            if (d == 1) {
                a.flags |= PF_KPOST;
                a.kmode = KM_IMAG;
                a.ck_post = ck;
            }
            if ((rc = launch_pot<R>(c, g0, a))) return rc;
            if (d > 1 && (rc = kinetic_block<R>(c, psi, KM_IMAG, ck, nullptr))) return rc;
        }
        // last potential half-step, then the Gram-Schmidt partial dots
        PassArgs<R> a = base_args<R>(c, psi, g0);
        a.scale1 = inv_n0;
        a.flags = PF_V1 | PF_SECOND;
        a.vmode = VM_IMAG;
        a.cv1 = c->as_scaled[2 * s - 1];
        if (gs && !split) {
            a.flags |= PF_GS_DOTS;
            a.prev = state_ptr<R>(c, 0);
            a.nprev = nprev;
            a.partials = c->d_part;
            a.npart = npart;
        }
        if ((rc = launch_pot<R>(c, g0, a))) return rc;
        if (gs) {
            ProfScope ps(c, PC_SMALL);
            if (split) gs_split<R>(c, nprev, true, psi, coef);
            const double *rows = c->d_part;
            int nrows = dot_rows;
            if (gs_blocks > 0) {
                // first-level sums of the rows in gs_blocks fixed slices (many
                // CUs), so that the one-block coefficient kernel sums a few
                // hundred rows instead of tens of thousands
                double *sums = c->d_part + (size_t)dot_rows * npart;
                hipLaunchKernelGGL(k_row_blocks, dim3(gs_blocks), dim3(256), 0, c->stream, c->d_part, dot_rows, npart,
                                   sums);
                rows = sums;
                nrows = gs_blocks;
            }
            hipLaunchKernelGGL(k_gs_coeffs, dim3(1), dim3(reduce_threads(nrows)), 0, c->stream, rows, nrows, npart,
                               nprev, c->P.dNr * inv_n0, c->d_small + SMALL_GRAM, coef, (double)it);
            HIPCK(c, hipGetLastError());
            pending = true;
            if ((it + 1) % GS_CHUNK == 0 && it + 1 < nsteps) {
                const int b = (int)(chunks_out & 1);
                HIPCK(c, hipMemcpyAsync(c->h_gs + 4 * b, coef + 2 * nprev + 1, 3 * sizeof(double),
                                        hipMemcpyDeviceToHost, c->stream));
                HIPCK(c, hipEventRecord(c->ev_gs[b], c->stream));
                if (chunks_out >= 1) {  // the previous chunk's copy; this chunk is already queued
                    HIPCK(c, hipEventSynchronize(c->ev_gs[1 - b]));
                    if (c->h_gs[4 * (1 - b)] != 0.0) {
                        HIPCK(c, hipStreamSynchronize(c->stream));
                        return collapse_error(c->h_gs + 4 * (1 - b));
                    }
                }
                ++chunks_out;
            }
        }
    }
    if (pending) {
        ProfScope ps(c, PC_SMALL);
        if (nprev <= QND_GRAM_MAX)
            gs_split<R>(c, nprev, false, psi, coef);
        else
            hipLaunchKernelGGL(k_elem<R>, dim3(flat_grid(c->N, 4, 1024, 16384)), dim3(256),
                               0, c->stream, (int)EM_GS_UPDATE, psi,
                               (const C<R> *)psi, 0, 0.0, c->N, c->P, (const C<R> *)state_ptr<R>(c, 0),
                               c->N, nprev, (const double *)coef);
        HIPCK(c, hipGetLastError());
    }
    if (gs) {
        double h[3] = {0, 0, 0};
        HIPCK(c, hipMemcpyAsync(h, coef + 2 * nprev + 1, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCK(c, hipStreamSynchronize(c->stream));
        if (h[0] != 0.0) return collapse_error(h);
    }
    prof_flush(c);
    return 0;
}

// standalone Gram-Schmidt in position space (no step)
template <typename R> int gram_schmidt(qnd_ctx *c, int slot)
{
    int rc;
    for (int j = 0; j <= slot; ++j)
        if ((rc = ensure_rep<R>(c, j, REP_R))) return rc;
    C<R> *v = state_ptr<R>(c, slot);
    // sequential modified Gram-Schmidt, as the reference
    std::vector<double> coef(2 * slot + 3, 0.0);
    for (int j = 0; j < slot; ++j) {
        double b[2];
        if ((rc = reduce<R>(c, RM_DOT, state_ptr<R>(c, j), v, nullptr, 0, 0, nullptr, 2, c->P.dNr, b)))
            return rc;
        std::vector<double> cj(3, 0.0);
        cj[0] = b[0];
        cj[1] = b[1];
        cj[2] = 1.0;
        HIPCK(c, hipMemcpyAsync(c->d_small + SMALL_COEF, cj.data(), 3 * sizeof(double),
                                hipMemcpyHostToDevice, c->stream));
        ProfScope ps(c, PC_SMALL);
        hipLaunchKernelGGL(k_elem<R>, dim3(flat_grid(c->N, 4, 1024, 16384)), dim3(256),
                           0, c->stream, (int)EM_GS_UPDATE, v,
                           (const C<R> *)v, 0, 0.0, c->N, c->P, (const C<R> *)state_ptr<R>(c, j),
                           c->N, 1, (const double *)(c->d_small + SMALL_COEF));
        HIPCK(c, hipGetLastError());
    }
    double nv[2];
    if ((rc = reduce<R>(c, RM_DOT, v, v, nullptr, 0, 0, nullptr, 2, c->P.dNr, nv))) return rc;
    if (nv[0] < 1.e-12)
        return fail(c, QND_EGS_COLLAPSE,
                    "ERROR: Wavefunction norm decayed to zero during Gram-Schmidt for state %d.", slot);
    if ((rc = elem<R>(c, EM_SCALE, v, v, 0, 1.0 / std::sqrt(nv[0])))) return rc;
    return 0;
}

// ---------------------------------------------------------------------
// real time (src/QuantumND_real.c:166-228)
// ---------------------------------------------------------------------
bool fused_codes(const int *codes, int ncodes)
{
    for (int k = 0; k < ncodes; ++k)
        if (!(codes[k] == QND_OBS_NORM || codes[k] == QND_OBS_DIPOLE_R || codes[k] == QND_OBS_DIPOLE_A ||
              codes[k] == QND_OBS_NONE))
            return false;
    return true;
}

template <typename R>
int real_steps_impl(qnd_ctx *c, int slot0, int nb, int gauge, int64_t step0, int64_t nsteps, const int *codes,
                    int ncodes, double *obs_out)
{
    const int d = c->dim, s = c->s;
    const bool batched = nb > 1;
    const int64_t ntime = batched ? c->bntime : c->ntime;
    int rc;
    // the tables are [ntime][s][dim] for the stage count they were set with
    if (ntime > 0 && (batched ? c->bfield_s : c->field_s) != s)
        return fail(c, QND_ESTATE, "the field tables were laid out for %d stages but the integrator has %d "
                    "(set the field again after qnd_set_integrator)", batched ? c->bfield_s : c->field_s, s);
    if (step0 < 0 || step0 + nsteps > ntime)
        return fail(c, QND_EINVAL, "steps [%lld, %lld) outside the field table of %lld steps",
                    (long long)step0, (long long)(step0 + nsteps), (long long)ntime);
    int width = 0;
    for (int k = 0; k < ncodes; ++k) width += obs_width(codes[k], d);
    const bool fused = fused_codes(codes, ncodes);
    const bool want_obs = width > 0;
    if (batched && want_obs && !fused)
        return fail(c, QND_EUNSUPPORTED, "batched runs record dipole_r, dipole_a and norm only");
    if (ncodes > 8) return fail(c, QND_EINVAL, "at most 8 observables per step");
    if (want_obs && fused) {
        // records of at least 4096 steps, so that the short calls after a
        // first one allocate nothing
        size_t need = (size_t)nb * nsteps * width;
        if (need > c->rec_cap) {
            need = std::max(need, (size_t)nb * 4096 * width);
            if (c->d_rec) HIPCK(c, hipFree(c->d_rec));
            c->d_rec = nullptr;
            c->rec_cap = 0;
            HIPCK(c, hipMalloc(&c->d_rec, need * sizeof(double)));
            c->rec_cap = need;
        }
    }
    // the slots of a batch move together: bring them to one state first
    for (int b = 1; b < nb; ++b) {
        const SlotState &s0 = c->slots[slot0], &sb = c->slots[slot0 + b];
        if (sb.rep != s0.rep || sb.pend_mask != s0.pend_mask || sb.pend_vend != s0.pend_vend) {
            for (int k = 0; k < nb; ++k)
                if ((rc = ensure_rep<R>(c, slot0 + k, REP_R))) return rc;
            break;
        }
    }
    // the length-gauge step starts from the plain state: a velocity-gauge V
    // half-step owed by earlier steps is paid now, at its own time
    if (gauge == QND_GAUGE_LENGTH && c->slots[slot0].pend_vend)
        for (int k = 0; k < nb; ++k)
            if ((rc = ensure_rep<R>(c, slot0 + k, REP_R))) return rc;
    SlotState &st = c->slots[slot0];
    C<R> *psi = state_ptr<R>(c, slot0);
    const Geo g0 = geometry(c, 0, nb);
    if ((rc = ensure_part(c, (size_t)g0.grid * QND_NOBS))) return rc;
    const double inv_n0 = 1.0 / (double)c->nr[0];
    ObsSpec spec;
    memset(&spec, 0, sizeof spec);
    spec.ncodes = ncodes;
    for (int k = 0; k < ncodes; ++k) spec.codes[k] = codes[k];
    // Merged step boundary (2-D, one instance, length gauge): between two
    // steps of this call, the last kinetic pass of step n saves the state
    // (the record's state, before the owed mask) to a ring of side buffers,
    // applies the axis-1 parts of the mask and of step n+1's first kinetic
    // factor, and the first axis-0 pass of step n+1 applies their axis-0
    // parts before its potential: the operators of the boundary pass and of
    // the first kinetic block, regrouped by axis (parts on different axes
    // commute).  The records of the saved states are computed by one batched
    // observable pass per ring fill, off the per-step chain.  Two passes per
    // step fewer.
    const int side_slots_max = 64;
    const size_t sbytes = (size_t)c->N * sizeof(C<R>);
    int side_slots = 0;
    bool merge = gauge == QND_GAUGE_LENGTH && d == 2 && !batched && (!want_obs || fused) && nsteps > 1;
    if (merge && want_obs) {
        side_slots = (int)std::min<int64_t>(side_slots_max, std::min<int64_t>(nsteps - 1,
                                            (int64_t)(QND_SIDE_MAX_BYTES / sbytes)));
        if (side_slots < 1) merge = false;
    }
    const typename TwT<R>::T *atab = nullptr;
    Geo gside{};
    if (merge) {
        atab = atab_for<R>(c, &rc);
        if (rc) return rc;
        if (side_slots > 0) {
            if (c->side_cap < (size_t)side_slots * sbytes) {
                // the ring's full size at once (later, longer calls reuse it)
                const size_t cap = (size_t)std::min<int64_t>(side_slots_max, (int64_t)(QND_SIDE_MAX_BYTES / sbytes));
                if (c->d_side) HIPCK(c, hipFree(c->d_side));
                c->d_side = nullptr;
                c->side_cap = 0;
                HIPCK(c, hipMalloc(&c->d_side, cap * sbytes));
                c->side_cap = cap * sbytes;
            }
            // the record pass's geometry for a full ring (its partial flushes
            // take the same tiles over fewer instances), so that calls of
            // different lengths reuse one launch shape
            gside = geometry(c, 0, (int)(c->side_cap / sbytes), -1, false);
            if ((rc = ensure_part(c, (size_t)std::max(g0.grid, gside.grid) * QND_NOBS))) return rc;
        }
    }
    C<R> *side = reinterpret_cast<C<R> *>(c->d_side);
    int side_n = 0;           // saved record states in the ring
    int64_t side_first = 0;   // step of the first of them
    // records of the saved states: one observable-only axis-0 pass over the
    // ring (a batch of side_n instances), then one finalize block per record
    auto flush_side = [&]() -> int {
        if (side_n == 0) return 0;
        Geo gs = gside;
        gs.nlines = (c->N / gs.n) * side_n;
        gs.grid = (int)(gs.nlines / gs.TW);
        PassArgs<R> a = base_args<R>(c, side, gs);
        a.scale1 = inv_n0;
        a.flags = PF_OBS | PF_MASK | PF_NOSTORE;
        a.partials = c->d_part;
        a.npart = QND_NOBS;
        a.B.nb = side_n;
        int r;
        if ((r = launch<R, PK_POTO, +1>(c, gs, a, PC_SMALL))) return r;
        Finalize f;
        memset(&f, 0, sizeof f);
        f.part = c->d_part;
        f.nwg = gs.grid / side_n;
        f.rec = c->d_rec + (side_first - step0) * width;
        f.spec = spec;
        f.B.nb = side_n;
        f.B.rec_stride = width;
        f.B.finE = c->d_Eo + side_first * d;
        f.B.finE_stride = d;
        if ((r = launch_finalize(c, &f))) return r;
        side_n = 0;
        return 0;
    };
    bool merged_in = false;   // this step's boundary was done by the previous step's passes
    // per-instance field rows: [nb][ntime][s][d] (stage) and [nb][ntime][d] (t_n)
    const int64_t stage_stride = ntime * s * d, obs_stride = ntime * d;
    auto stage_rows = [&](const double *tab, int64_t it, int ia) {
        FieldRows fr;
        if (batched) {
            fr.base = tab + (it * s + ia) * d;
            fr.stride = stage_stride;
        }
        return fr;
    };

    for (int64_t it = step0; it < step0 + nsteps; ++it) {
        const double *Et = batched ? nullptr : &c->Et[(size_t)it * s * d];
        const double *At = batched ? nullptr : &c->At[(size_t)it * s * d];
        const double *Eo = batched ? nullptr : &c->Eobs[(size_t)it * d];
        const double *Ao = batched ? nullptr : &c->Aobs[(size_t)it * d];
        // --- observables at the step boundary (unfused codes: position-space evaluation) ---
        if (want_obs && !fused) {
            if ((rc = observables<R>(c, slot0, codes, ncodes, gauge == QND_GAUGE_VELOCITY ? Ao : nullptr, Eo,
                                     obs_out + (it - step0) * width)))
                return rc;
        }
        // a slot in position space owes nothing; the boundary pass pays the
        // owed mask / half-step of a slot in the axis-0 Fourier representation
        if (st.rep == REP_R) {
            if ((rc = plain_pass<R>(c, psi, 0, -1, 1.0, nb))) return rc;
            st.rep = REP_M;
        }
        Finalize fin;
        memset(&fin, 0, sizeof fin);
        const Finalize *finp = nullptr;
        if (want_obs && fused && !merged_in) {
            if (!batched)
                for (int j = 0; j < d; ++j) spec.E[j] = Eo[j];
            fin.part = c->d_part;
            fin.nwg = g0.grid / nb;
            fin.rec = c->d_rec + (it - step0) * width;
            fin.spec = spec;
            fin.B.nb = nb;
            fin.B.rec_stride = nsteps * width;
            if (batched) {
                fin.B.finE = c->d_bEo + it * d;
                fin.B.finE_stride = obs_stride;
            }
            finp = &fin;
        }
        auto finalize_now = [&]() -> int {
            if (!finp) return 0;
            int r = launch_finalize(c, finp);
            finp = nullptr;
            return r;
        };
        auto pot_args = [&]() {
            PassArgs<R> a = base_args<R>(c, psi, g0);
            a.scale1 = inv_n0;
            a.B.nb = nb;
            return a;
        };
        auto set_rows = [&](PassArgs<R> &a, FieldRows fr) {
            a.B.fld = fr.base;
            a.B.fld_stride = fr.stride;
        };
        if (gauge == QND_GAUGE_LENGTH) {
            if (!merged_in) {
                // boundary pass: owed mask, observables
                PassArgs<R> a = pot_args();
                a.flags = PF_SECOND;
                if (st.pend_mask) a.flags |= PF_MASK;
                if (finp) {
                    a.flags |= PF_OBS;
                    a.partials = c->d_part;
                    a.npart = QND_NOBS;
                }
                if (a.flags != PF_SECOND) {
                    if ((rc = launch_pot<R>(c, g0, a))) return rc;
                }
                st.pend_mask = false;
                if (d == 1 && (rc = finalize_now())) return rc;
            }
            bool merged_out = false;
            for (int ia = 0; ia < s; ++ia) {
                const double a0 = c->as_scaled[2 * ia], a1 = c->as_scaled[2 * ia + 1];
                if (d > 1 && ia == 0 && !merged_in) {
                    if ((rc = kinetic_block<R>(c, psi, KM_PHASE, a0, nullptr, finp, nb))) return rc;
                    finp = nullptr;
                }
                PassArgs<R> b = pot_args();
                b.flags = PF_V1 | PF_SECOND;
                b.vmode = VM_LENGTH;
                b.cv1 = a0 + a1;
                if (batched) set_rows(b, stage_rows(c->d_bEt, it, ia));
                else
                    for (int j = 0; j < d; ++j) b.Ev[j] = Et[ia * d + j];
                if (d == 1) {
                    b.flags |= PF_KPRE | PF_KPOST;
                    b.kmode = KM_PHASE;
                    b.ck_pre = a0;
                    b.ck_post = a1;
                }
                if (ia == 0 && merged_in) {
                    // the owed axis-0 mask and axis-0 part of K(a_0)
                    b.flags |= PF_PREKX | (c->P.mask_on ? PF_MASK : 0u);
                    b.atab = atab;
                }
                if ((rc = launch_pot<R>(c, g0, b))) return rc;
                if (d > 1) {
                    double ck = a1 + (ia < s - 1 ? c->as_scaled[2 * ia + 2] : 0.0);
                    if (ia == s - 1 && merge && it + 1 < step0 + nsteps) {
                        // merged boundary: K(a_last), save the record's state,
                        // axis-1 parts of the mask and of K(a_0)
                        const Geo gk = geometry(c, 1, 1, PK_KIN);
                        PassArgs<R> k = base_args<R>(c, psi, gk);
                        k.flags = c->P.mask_on ? PF_MASK : 0u;
                        k.kmode = KM_PHASE;
                        k.ck_mid = ck;
                        k.scale2 = 1.0 / (double)c->nr[1];
                        k.ktab = ktab_for<R>(c, ck, gk, &rc);
                        if (rc) return rc;
                        if (k.ktab) k.flags |= PF_KTAB;
                        k.atab = atab;
                        if (want_obs) {
                            if (side_n == 0) side_first = it + 1;
                            k.flags |= PF_SAVE;
                            k.save = side + (size_t)side_n * c->N;
                            ++side_n;
                        }
                        if ((rc = launch<R, PK_KINS, -1>(c, gk, k, PC_BND))) return rc;
                        merged_out = true;
                    } else if ((rc = kinetic_block<R>(c, psi, KM_PHASE, ck, nullptr, nullptr, nb))) {
                        return rc;
                    }
                }
            }
            st.pend_mask = c->P.mask_on != 0;
            merged_in = merged_out;
            if (side_n == side_slots && (rc = flush_side())) return rc;
        } else {
            // velocity gauge: boundary pass carries the owed last half-step,
            // the owed mask, the observables and the first half-step
            for (int ia = 0; ia < s; ++ia) {
                PassArgs<R> a = pot_args();
                a.flags = PF_V2 | PF_SECOND;
                a.vmode = VM_VELOCITY;
                a.cv2 = c->as_scaled[2 * ia];
                if (ia == 0) {
                    if (st.pend_vend) {
                        a.flags |= PF_V1;
                        a.cv1 = c->as_scaled[2 * s - 1];
                    }
                    if (st.pend_mask) a.flags |= PF_MASK;
                    if (want_obs && fused) {
                        a.flags |= PF_OBS;
                        a.partials = c->d_part;
                        a.npart = QND_NOBS;
                    }
                } else {
                    a.flags |= PF_V1;
                    a.cv1 = c->as_scaled[2 * ia - 1];
                }
                const double ck = c->as_scaled[2 * ia] + c->as_scaled[2 * ia + 1];
                if (d == 1) {
                    a.flags |= PF_KPOST;
                    a.kmode = KM_VELOCITY;
                    a.ck_post = ck;
                    if (batched) set_rows(a, stage_rows(c->d_bAt, it, ia));
                    else a.Ak[0] = At[ia * d];
                }
                if ((rc = launch_pot<R>(c, g0, a))) return rc;
                if (ia == 0) {
                    st.pend_vend = st.pend_mask = false;
                    if (d == 1 && (rc = finalize_now())) return rc;
                }
                if (d > 1) {
                    if ((rc = kinetic_block<R>(c, psi, KM_VELOCITY, ck, batched ? nullptr : &At[ia * d],
                                               ia == 0 ? finp : nullptr, nb,
                                               batched ? stage_rows(c->d_bAt, it, ia) : FieldRows())))
                        return rc;
                    if (ia == 0) finp = nullptr;
                }
            }
            st.pend_vend = true;
            st.pend_mask = c->P.mask_on != 0;
        }
    }
    if ((rc = flush_side())) return rc;
    for (int b = 1; b < nb; ++b) c->slots[slot0 + b] = st;
    if (want_obs && fused) {
        // device records [nb][nsteps][width] -> obs_out [nb][nsteps][width]
        HIPCK(c, hipMemcpyAsync(obs_out, c->d_rec, (size_t)nb * nsteps * width * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream));
    }
    HIPCK(c, hipStreamSynchronize(c->stream));
    prof_flush(c);
    return 0;
}

template <typename R>
int real_steps(qnd_ctx *c, int slot, int gauge, int64_t step0, int64_t nsteps, const int *codes, int ncodes,
               double *obs_out)
{
    return real_steps_impl<R>(c, slot, 1, gauge, step0, nsteps, codes, ncodes, obs_out);
}

// ---------------------------------------------------------------------
// Adjoint of the length-gauge propagation (dim >= 2)
//
// Forward step n, as real_steps runs it on the axis-0 Fourier
// representation M:  B_n, K(k_0), P_0, K(k_1), ..., P_{s-1}, K(k_s), with
//   B_n = F0 Mask F0^-1         (the owed mask; observables of y_n = Mask F0^-1 M)
//   P_i = F0 Phi_i F0^-1,       Phi_i = exp(-c_i Vcap - i c_i (V + x.E_{n,i}))
//   K(k) = the all-k kinetic factor exp(-i k Kin) between axis-1.. transforms.
// Every factor's adjoint has its own form with the phase conjugated
// (F0^dagger = n0 F0^-1), so the costate runs backwards through the same
// passes with PF_CONJ.  With g = dL/dM* (dL = 2 Re <g, dM>):
//   dL/dE_{n,i,j}  = 2 n0 c_i sum_r x_j Im(conj(u) phi)      u = F0^-1 g after P_i,
//                                                           phi = Phi_i F0^-1 M (saved)
//   B_n^dagger g + source = F0 Mask (F0^-1 g + sum_j s_j dNr/n0 (w_j - E_j) y_n)
// for the dipole acceleration a_{n,j} = dNr sum (w_j - E_j)|y_n|^2, w = -grad V,
// s = dL/da; the final norm enters as a source dNr g_norm / n0 * y_N.
// Forward states are checkpointed every K steps; each segment is recomputed
// once, saving y_n and phi_{n,i} in position space (s+1 states per step).
// ---------------------------------------------------------------------
template <typename R> int adj_forward(qnd_ctx *c, int slot, int64_t nsteps, int64_t K, double *obs_out,
                                      double *norm_final)
{
    const int d = c->dim;
    int rc;
    const int64_t nseg = (nsteps + K - 1) / K;
    const size_t sbytes = (size_t)c->N * sizeof(C<R>);
    if (c->adj_ckpt_cap < (size_t)nseg) {
        if (c->d_adj_ckpt) HIPCK(c, hipFree(c->d_adj_ckpt));
        c->d_adj_ckpt = nullptr;
        HIPCK(c, hipMalloc(&c->d_adj_ckpt, sbytes * nseg));
        c->adj_ckpt_cap = nseg;
    }
    if (!c->d_adj_final) HIPCK(c, hipMalloc(&c->d_adj_final, sbytes));
    c->adj_ckpt_state.assign(nseg, SlotState());
    c->adj_obs.assign((size_t)nsteps * (d + 1), 0.0);
    const int codes[2] = {QND_OBS_DIPOLE_A, QND_OBS_NORM};
    C<R> *psi = state_ptr<R>(c, slot);
    for (int64_t sg = 0; sg < nseg; ++sg) {
        const int64_t n0 = sg * K, len = (n0 + K <= nsteps) ? K : nsteps - n0;
        HIPCK(c, hipMemcpyAsync(reinterpret_cast<C<R> *>(c->d_adj_ckpt) + sg * c->N, psi, sbytes,
                                hipMemcpyDeviceToDevice, c->stream));
        c->adj_ckpt_state[sg] = c->slots[slot];
        if ((rc = real_steps<R>(c, slot, QND_GAUGE_LENGTH, n0, len, codes, 2, c->adj_obs.data() + n0 * (d + 1))))
            return rc;
    }
    // the final boundary: owed mask, then the norm of y_N (kept for the costate)
    double nf = 0.0;
    const int cn = QND_OBS_NORM;
    if ((rc = observables<R>(c, slot, &cn, 1, nullptr, nullptr, &nf))) return rc;
    HIPCK(c, hipMemcpyAsync(c->d_adj_final, psi, sbytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    c->adj_slot = slot;
    c->adj_nsteps = nsteps;
    c->adj_K = K;
    c->adj_prec = c->prec;
    c->adj_gen = c->setup_gen;
    if (obs_out) memcpy(obs_out, c->adj_obs.data(), c->adj_obs.size() * sizeof(double));
    if (norm_final) *norm_final = nf;
    return 0;
}

template <typename R>
int adj_backward(qnd_ctx *c, const double *dL_da, double dL_dnorm, double *dL_dEt, double *dL_dEobs, double *dL_dsoft)
{
    const int d = c->dim, s = c->s;
    const int64_t nsteps = c->adj_nsteps, K = c->adj_K, nseg = (nsteps + K - 1) / K;
    const double n0d = (double)c->nr[0], dNr = c->P.dNr, inv_n0 = 1.0 / n0d;
    int rc;
    const size_t sbytes = (size_t)c->N * sizeof(C<R>);
    const Geo g0 = geometry(c, 0);
    const int nwg = g0.grid, per_step = s + 1;
    // double-buffered saves and partials: the recompute of segment k-1 (second
    // stream) overlaps the adjoint sweep of segment k (first stream)
    const size_t nsave = (size_t)K * per_step;
    if (c->adj_save_cap < 2 * nsave) {
        if (c->d_adj_save) HIPCK(c, hipFree(c->d_adj_save));
        c->d_adj_save = nullptr;
        HIPCK(c, hipMalloc(&c->d_adj_save, sbytes * 2 * nsave));
        c->adj_save_cap = 2 * nsave;
    }
    // partials of the sweep's passes, their sums; the boundary source pass's
    // partials and sums; the per-step source weights
    const int64_t lpi0 = c->N / c->nr[0];
    const size_t npart = 2 * nsave * nwg * QND_NOBS, ngsum = (size_t)nsteps * per_step * QND_NOBS;
    const size_t npartT = (size_t)K * lpi0 * QND_NOBS, ngsumT = (size_t)nsteps * QND_NOBS, nsrc = (size_t)nsteps * d;
    const size_t nadj = npart + ngsum + npartT + ngsumT + nsrc;
    if (c->adj_part_cap < nadj) {
        if (c->d_adj_part) HIPCK(c, hipFree(c->d_adj_part));
        c->d_adj_part = nullptr;
        HIPCK(c, hipMalloc(&c->d_adj_part, nadj * sizeof(double)));
        c->adj_part_cap = nadj;
    }
    double *gsum_dev = c->d_adj_part + npart;
    double *partT = gsum_dev + ngsum, *gsumT_dev = partT + npartT, *src_dev = gsumT_dev + ngsumT;
    {
        // source weights of the records, dL/da_{n,j} dNr / n0, per step
        std::vector<double> sw(nsrc);
        for (size_t q = 0; q < nsrc; ++q) sw[q] = dL_da[q] * dNr * inv_n0;
        HIPCK(c, hipMemcpy(src_dev, sw.data(), nsrc * sizeof(double), hipMemcpyHostToDevice));
    }
    C<R> *lam = reinterpret_cast<C<R> *>(c->d_scr[0]);
    C<R> *W = reinterpret_cast<C<R> *>(c->d_scr[1]);
    // a segment's saves: the record states y_k (or, at merged boundaries, their
    // axis-0 Fourier representation) in K contiguous slots, then phi_{k,i}
    auto saves_of = [&](int64_t sg) { return reinterpret_cast<C<R> *>(c->d_adj_save) + (sg & 1) * nsave * c->N; };
    auto yslot = [&](C<R> *saves, int64_t k) { return saves + (size_t)k * c->N; };
    auto pslot = [&](C<R> *saves, int64_t k, int ia) { return saves + (size_t)(K + k * s + ia) * c->N; };
    auto part_of = [&](int64_t sg) { return c->d_adj_part + (sg & 1) * nsave * nwg * QND_NOBS; };
    std::vector<hipEvent_t> ev_rec(nseg, nullptr), ev_adj(nseg, nullptr);
    for (int64_t k = 0; k < nseg; ++k) {
        HIPCK(c, hipEventCreateWithFlags(&ev_rec[k], hipEventDisableTiming));
        HIPCK(c, hipEventCreateWithFlags(&ev_adj[k], hipEventDisableTiming));
    }
    auto cleanup = [&]() {
        for (auto e : ev_rec)
            if (e) hipEventDestroy(e);
        for (auto e : ev_adj)
            if (e) hipEventDestroy(e);
    };
    std::vector<std::vector<char>> masked(nseg);
    const double *as = c->as_scaled;
    auto kcoef = [&](int i) { return i == 0 ? as[0] : (i == s ? as[2 * s - 1] : as[2 * i - 1] + as[2 * i]); };
    auto pot = [&](C<R> *data) {
        PassArgs<R> a = base_args<R>(c, data, g0);
        a.scale1 = inv_n0;
        a.flags = PF_SECOND;
        return a;
    };
    auto kin_adj = [&](C<R> *data, double ck) -> int {
        return kinetic_block<R>(c, data, KM_PHASE, ck, nullptr, nullptr, 1, FieldRows(), PF_CONJ);
    };
    // Recompute of segment sg from its checkpoint on the second stream
    // (saving y_n and phi_{n,i}; it may start once the sweep that last read
    // its buffer is done), and the costate sweep of segment sg on the first
    // stream (once the recompute of sg is done).  The recompute of sg - 1 and
    // the sweep of sg are enqueued step by step, interleaved, so that both
    // streams have work queued together and run concurrently.
    // The recompute merges its step boundaries as real_steps does (the first
    // step of a segment starts from the checkpoint with the boundary pass):
    // the last kinetic pass of step k saves the record's state y_{k+1} in
    // the axis-0 Fourier representation, which the sweep's source pass turns
    // into position space itself (ysrcm).
    std::vector<std::vector<char>> ysrcm(nseg);
    const typename TwT<R>::T *atab = atab_for<R>(c, &rc);
    if (rc) return cleanup(), rc;
    struct Seg {
        int64_t sg = -1, nA = 0, len = 0, k = 0;
        C<R> *saves = nullptr;
        double *part = nullptr;
        SlotState st;
        bool merged_in = false;
    };
    auto seg_of = [&](int64_t sg) {
        Seg g;
        g.sg = sg;
        g.nA = sg * K;
        g.len = (g.nA + K <= nsteps) ? K : nsteps - g.nA;
        g.saves = saves_of(sg);
        g.part = part_of(sg);
        g.st = c->adj_ckpt_state[sg];
        return g;
    };
    struct OnStream2 {
        qnd_ctx *c;
        explicit OnStream2(qnd_ctx *cc) : c(cc) { std::swap(c->stream, c->stream2); }
        ~OnStream2() { std::swap(c->stream, c->stream2); }
    };
    auto rec_begin = [&](Seg &g) -> int {
        OnStream2 on2(c);
        if (g.sg + 2 < nseg) HIPCK(c, hipStreamWaitEvent(c->stream, ev_adj[g.sg + 2], 0));
        HIPCK(c, hipMemcpyAsync(W, reinterpret_cast<C<R> *>(c->d_adj_ckpt) + g.sg * c->N, sbytes,
                                hipMemcpyDeviceToDevice, c->stream));
        if (g.st.rep == REP_R) {
            if (int r = plain_pass<R>(c, W, 0, -1, 1.0)) return r;
            g.st.rep = REP_M;
        }
        masked[g.sg].assign(g.len, 0);
        ysrcm[g.sg].assign(g.len, 0);
        return 0;
    };
    auto rec_step = [&](Seg &g) -> int {
        OnStream2 on2(c);
        const int64_t k = g.k++, n = g.nA + k;
        const double *Et = &c->Et[(size_t)n * s * d];
        masked[g.sg][k] = g.st.pend_mask;
        const bool merged_in = g.merged_in;
        if (!merged_in) {
            PassArgs<R> b = pot(W);
            b.flags |= PF_SAVE | (g.st.pend_mask ? PF_MASK : 0u);
            b.save = yslot(g.saves, k);
            if (int r = launch<R, PK_POT, +1>(c, g0, b, PC_POT)) return r;
            if (int r = kinetic_block<R>(c, W, KM_PHASE, kcoef(0), nullptr)) return r;
        }
        g.merged_in = false;
        for (int ia = 0; ia < s; ++ia) {
            PassArgs<R> p = pot(W);
            p.flags |= PF_V1 | PF_SAVE;
            p.vmode = VM_LENGTH;
            p.cv1 = as[2 * ia] + as[2 * ia + 1];
            for (int j = 0; j < d; ++j) p.Ev[j] = Et[ia * d + j];
            p.save = pslot(g.saves, k, ia);
            if (ia == 0 && merged_in) {
                p.flags |= PF_PREKX | (g.st.pend_mask ? PF_MASK : 0u);
                p.atab = atab;
                if (int r = launch<R, PK_POT, +1>(c, g0, p, PC_POT)) return r;
            } else if (int r = launch<R, PK_POTS, +1>(c, g0, p, PC_POT)) {
                return r;
            }
            if (ia == s - 1 && k + 1 < g.len && d == 2) {
                // merged boundary: K(a_last), save y_{k+1} (axis-0 Fourier), axis-1 parts
                const Geo gk = geometry(c, 1, 1, PK_KIN);
                PassArgs<R> q = base_args<R>(c, W, gk);
                q.flags = PF_SAVE | (c->P.mask_on ? PF_MASK : 0u);
                q.kmode = KM_PHASE;
                q.ck_mid = kcoef(s);
                q.scale2 = 1.0 / (double)c->nr[1];
                int r = 0;
                q.ktab = ktab_for<R>(c, q.ck_mid, gk, &r);
                if (r) return r;
                if (q.ktab) q.flags |= PF_KTAB;
                q.atab = atab;
                q.save = yslot(g.saves, k + 1);
                if ((r = launch<R, PK_KINS, -1>(c, gk, q, PC_KIN))) return r;
                ysrcm[g.sg][k + 1] = 1;
                g.merged_in = true;
            } else if (int r = kinetic_block<R>(c, W, KM_PHASE, kcoef(ia + 1), nullptr)) {
                return r;
            }
        }
        g.st.pend_mask = c->P.mask_on != 0;
        return 0;
    };
    auto rec_end = [&](Seg &g) -> int {
        OnStream2 on2(c);
        HIPCK(c, hipEventRecord(ev_rec[g.sg], c->stream));
        return 0;
    };
    // The sweep crosses a merged boundary (records saved in the axis-0 Fourier
    // representation, 2-D) in two passes, the adjoints of the forward's two
    // boundary passes: the adjoint B_k^dagger = F0 Mask F0^-1 + T_k of the
    // record's pass is linear in the costate plus the source
    // T_k = F0 Mask S_k(y_k), which does not depend on the costate.  Before the
    // segment's steps, one batched pass turns the saved states of its merged
    // steps into T_k in place (PF_ZERO: source alone), with the explicit
    // soft-core partials of those records.  Then
    //   pass A (axis 0): P_0^dagger, the conjugated axis-0 part of K(a_0), the
    //                    axis-0 mask (PF_POSTKX);
    //   pass B (axis 1): the conjugated axis-1 part of K(a_0), the axis-1 mask,
    //                    + T_k, then K(a_last)^dagger of step k-1 (PK_KINR).
    auto sweep_begin = [&](Seg &g) -> int {
        HIPCK(c, hipStreamWaitEvent(c->stream, ev_rec[g.sg], 0));
        g.k = g.len - 1;
        int64_t k1 = 1;
        while (k1 < g.len && !ysrcm[g.sg][k1]) ++k1;
        const int64_t ni = g.len - k1;  // merged steps k1 .. len-1
        for (int64_t k = k1; k < g.len; ++k)
            if (!ysrcm[g.sg][k] || masked[g.sg][k] != masked[g.sg][k1]) return fail(c, QND_EINVAL, "adjoint: mixed boundaries");
        if (ni > 0) {
            const Geo gt = geometry(c, 0, (int)ni);
            PassArgs<R> a = base_args<R>(c, yslot(g.saves, k1), gt);
            a.scale1 = inv_n0;
            a.flags = PF_SRC | PF_SRCM | PF_ZERO | PF_SECOND | (masked[g.sg][k1] ? PF_MASK : 0u);
            a.save = yslot(g.saves, k1);
            a.partials = partT;
            a.npart = QND_NOBS;
            a.B.nb = (int)ni;
            a.B.fld = c->d_Eo + (g.nA + k1) * d;
            a.B.fld_stride = d;
            a.B.srcr = src_dev + (g.nA + k1) * d;
            a.B.srcr_stride = d;
            if (int r = launch<R, PK_POT, +1>(c, gt, a, PC_POT)) return r;
            ProfScope ps(c, PC_SMALL);
            hipLaunchKernelGGL(k_sum_blocks, dim3(((int)ni * QND_NOBS + 3) / 4), dim3(256), 0, c->stream,
                               (const double *)partT, (int)ni, (int)(gt.grid / ni), QND_NOBS,
                               gsumT_dev + (g.nA + k1) * QND_NOBS);
            HIPCK(c, hipGetLastError());
        }
        return 0;
    };
    auto sweep_step = [&](Seg &g) -> int {
        const int64_t k = g.k--, n = g.nA + k;
        const double *Et = &c->Et[(size_t)n * s * d];
        const double *Eo = &c->Eobs[(size_t)n * d];
        // K(a_last)^dagger of this step, unless the merged boundary after it
        // (the next step's pass B) applied it
        if (k == g.len - 1 || !ysrcm[g.sg][k + 1])
            if (int r = kin_adj(lam, kcoef(s))) return r;
        const bool merged = ysrcm[g.sg][k] != 0;
        for (int ia = s - 1; ia >= 0; --ia) {
            PassArgs<R> p = pot(lam);
            p.flags |= PF_V1 | PF_CONJ | PF_GRAD;
            p.vmode = VM_LENGTH;
            p.cv1 = as[2 * ia] + as[2 * ia + 1];
            for (int j = 0; j < d; ++j) p.Ev[j] = Et[ia * d + j];
            p.save = pslot(g.saves, k, ia);
            p.partials = g.part + (size_t)(k * per_step + 1 + ia) * nwg * QND_NOBS;
            p.npart = QND_NOBS;
            if (ia == 0 && merged) {
                p.flags |= PF_POSTKX | (masked[g.sg][k] ? PF_MASK : 0u);
                p.atab = atab;
                if (int r = launch<R, PK_POT, +1>(c, g0, p, PC_POT)) return r;
            } else if (int r = launch<R, PK_POTG, +1>(c, g0, p, PC_POT)) {
                return r;
            }
            if (ia > 0 || !merged)
                if (int r = kin_adj(lam, kcoef(ia))) return r;
        }
        if (merged) {
            // pass B: axis-1 parts of K(a_0) and of the mask, + T_k, K(a_last)^dagger of step k-1
            const Geo gk = geometry(c, 1, 1, PK_KIN);
            PassArgs<R> q = base_args<R>(c, lam, gk);
            q.flags = PF_CONJ | (masked[g.sg][k] ? PF_MASK : 0u);
            q.kmode = KM_PHASE;
            q.ck_mid = kcoef(s);
            q.scale2 = 1.0 / (double)c->nr[1];
            int r = 0;
            q.ktab = ktab_for<R>(c, q.ck_mid, gk, &r);
            if (r) return r;
            if (q.ktab) q.flags |= PF_KTAB;
            q.atab = atab;
            q.save = yslot(g.saves, k);
            return launch<R, PK_KINR, -1>(c, gk, q, PC_KIN);
        }
        PassArgs<R> b = pot(lam);
        b.flags |= PF_SRC | (masked[g.sg][k] ? PF_MASK : 0u);
        for (int j = 0; j < d; ++j) {
            b.src[j] = dL_da[n * d + j] * dNr * inv_n0;
            b.Ev[j] = Eo[j];
        }
        b.save = yslot(g.saves, k);
        b.partials = g.part + (size_t)(k * per_step) * nwg * QND_NOBS;
        b.npart = QND_NOBS;
        return launch<R, PK_POT, +1>(c, g0, b, PC_POT);
    };
    auto sweep_end = [&](Seg &g) -> int {
        {
            ProfScope ps(c, PC_SMALL);
            const int npass = (int)(g.len * per_step);
            hipLaunchKernelGGL(k_sum_blocks, dim3((npass * QND_NOBS + 3) / 4), dim3(256), 0, c->stream,
                               (const double *)g.part, npass, nwg, QND_NOBS, gsum_dev + g.nA * per_step * QND_NOBS);
            HIPCK(c, hipGetLastError());
        }
        HIPCK(c, hipEventRecord(ev_adj[g.sg], c->stream));
        return 0;
    };

    // costate after the final boundary: F0 Mask (dNr g_norm / n0 y_N)
    HIPCK(c, hipMemsetAsync(lam, 0, sbytes, c->stream));
    {
        PassArgs<R> a = pot(lam);
        a.flags |= PF_SRC | (c->P.mask_on ? PF_MASK : 0u);
        a.save = reinterpret_cast<C<R> *>(c->d_adj_final);
        a.src_norm = dL_dnorm * dNr * inv_n0;
        a.partials = c->d_adj_part;
        a.npart = QND_NOBS;
        if ((rc = launch<R, PK_POT, +1>(c, g0, a, PC_POT))) return cleanup(), rc;
    }
    // the second stream must not start before the checkpoints are written
    HIPCK(c, hipStreamSynchronize(c->stream));
    {
        Seg last = seg_of(nseg - 1);
        if ((rc = rec_begin(last))) return cleanup(), rc;
        while (last.k < last.len)
            if ((rc = rec_step(last))) return cleanup(), rc;
        if ((rc = rec_end(last))) return cleanup(), rc;
    }
    for (int64_t sg = nseg - 1; sg >= 0; --sg) {
        Seg sw = seg_of(sg), rc_seg;
        const bool rec = sg > 0;
        if (rec) {
            rc_seg = seg_of(sg - 1);
            if ((rc = rec_begin(rc_seg))) return cleanup(), rc;
        }
        if ((rc = sweep_begin(sw))) return cleanup(), rc;
        while (sw.k >= 0 || (rec && rc_seg.k < rc_seg.len)) {
            if (rec && rc_seg.k < rc_seg.len && (rc = rec_step(rc_seg))) return cleanup(), rc;
            if (sw.k >= 0 && (rc = sweep_step(sw))) return cleanup(), rc;
        }
        if (rec && (rc = rec_end(rc_seg))) return cleanup(), rc;
        if ((rc = sweep_end(sw))) return cleanup(), rc;
    }
    std::vector<double> gsum(ngsum), gsumT(ngsumT);
    HIPCK(c, hipMemcpyAsync(gsum.data(), gsum_dev, ngsum * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipMemcpyAsync(gsumT.data(), gsumT_dev, ngsumT * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream2));
    cleanup();
    std::vector<char> merged_rec(nsteps, 0);
    for (int64_t sg = 0; sg < nseg; ++sg)
        for (size_t k = 0; k < ysrcm[sg].size(); ++k) merged_rec[sg * K + k] = ysrcm[sg][k];
    for (int64_t k = 0; k < nsteps * s * d; ++k) dL_dEt[k] = 0.0;
    double dsoft = 0.0;
    for (int64_t n = 0; n < nsteps; ++n) {
        // the record's explicit partials: from its boundary pass, or from the
        // source pass of a merged boundary
        const double *bsum = merged_rec[n] ? &gsumT[(size_t)n * QND_NOBS] : &gsum[(size_t)(n * per_step) * QND_NOBS];
        dsoft += n0d * bsum[0];  // explicit dependence of a_n on the soft-core length
        const double nrm = c->adj_obs[n * (d + 1) + d];
        for (int j = 0; j < d; ++j) dL_dEobs[n * d + j] = -dL_da[n * d + j] * nrm;
        for (int ia = 0; ia < s; ++ia) {
            const double *ps = &gsum[(size_t)(n * per_step + 1 + ia) * QND_NOBS];
            const double cc = as[2 * ia] + as[2 * ia + 1];
            dsoft += 2.0 * n0d * cc * ps[0];
            for (int j = 0; j < d; ++j) dL_dEt[(n * s + ia) * d + j] = 2.0 * n0d * cc * ps[1 + j];
        }
    }
    if (dL_dsoft) *dL_dsoft = dsoft;
    prof_flush(c);
    return 0;
}

// ---------------------------------------------------------------------
// Several grids in one launch per pass (qnd_group_real_steps): the batched
// length-gauge step of each member context, every pass launched once over
// all of them (k_pass_group), the step sequence replayable through the
// device step counter, optionally as a captured graph.
// ---------------------------------------------------------------------
struct GroupGeo {
    std::vector<Geo> g;  // per member
    int threads = 0;
    size_t lds = 0;
};

// Launch geometry of a pass along `axis` for every member with one
// workgroup size: each member's own geometry, then its lines per workgroup
// rescaled to the largest member's thread count (halving it until every
// member admits it).  False when no common size fits.
bool group_geometry(qnd_ctx *const *cs, int n, int axis, int kind, GroupGeo &G)
{
    G.g.assign(n, Geo{});
    int want = 64, tmax = 1;  // at least a wavefront; at least every member's line
    for (int k = 0; k < n; ++k) {
        G.g[k] = geometry(cs[k], axis, cs[k]->bnb, kind);
        want = std::max(want, G.g[k].threads);
        tmax = std::max(tmax, G.g[k].T);
    }
    want = pow2floor(2 * want - 1);  // the power of two at or above
    for (int thr = std::min(want, QND_PASS_THREADS); thr >= tmax; thr /= 2) {
        bool ok = true;
        std::vector<Geo> out(G.g);
        for (int k = 0; k < n && ok; ++k) {
            Geo &g = out[k];
            qnd_ctx *c = cs[k];
            if (thr % g.T != 0) { ok = false; break; }
            const int tw = thr / g.T;
            const int64_t lpi = c->N / g.n;
            if ((tw & (tw - 1)) != 0 || lpi % tw != 0 || (int64_t)tw > g.nlines || (g.S > 1 && (g.S % tw) != 0) ||
                2 * (size_t)tw * (g.n + (tw > 1 ? 1 : 0)) * c->lds_esz > c->lds_max) {
                ok = false;
                break;
            }
            g.TW = tw;
            g.log2TW = ilog2(tw);
            g.threads = g.T * tw;
            g.grid = (int)(g.nlines / tw);
            g.ldsStride = g.n + (tw > 1 ? 1 : 0);
            g.lds = 2 * (size_t)tw * g.ldsStride * c->lds_esz;
        }
        if (ok) {
            G.g = out;
            G.threads = thr;
            G.lds = 0;
            for (auto &g : G.g) G.lds = std::max(G.lds, g.lds);
            return true;
        }
    }
    return false;
}

// One recorded launch of a group step: a pass kernel with its descriptor
// (an index into the call's descriptor array), or the step counter advance.
struct GLaunch {
    const void *fn;
    int grid, threads;
    size_t lds;
    int desc;  // -1: k_step_advance
};

// One member's pass in the branch form of a group step: the member's own
// pass kernel (its values per thread and geometry) with its arguments.
template <typename R> struct BLaunch {
    const void *fn;
    int grid, threads, nblk;
    size_t lds;
    unsigned geo, fl;
    PassArgs<R> a;
};
template <typename R> using BLists = std::vector<std::vector<BLaunch<R>>>;

// Record a pass over every member: its descriptor goes to `descs` (uploaded
// once per call), its launch to `out`; or, in the branch form (br given),
// each member's own launch to its list (*br)[k].
template <typename R, int KIND, int DIR>
void record_group(const GroupGeo &G, const std::vector<PassArgs<R>> &as, std::vector<GroupArgs<R>> &descs,
                  std::vector<GLaunch> &out, BLists<R> *br = nullptr)
{
    const int n = (int)as.size();
    if (br) {
        for (int k = 0; k < n; ++k) {
            const Geo &g = G.g[k];
            BLaunch<R> b;
            b.fn = pass_fn<R, KIND, DIR>(g.vpt);  // (x64 members are refused by the group call)
            const int extra = (KIND == PK_KIN && (as[k].flags & PF_FINALIZE)) ? as[k].B.nb : 0;
            b.grid = g.grid + extra;
            b.threads = g.threads;
            b.nblk = g.grid;
            b.lds = g.lds;
            b.geo = pass_geo(g.log2n, g.log2T, g.log2TW, as[k].P.pow2);
            b.fl = as[k].flags;
            b.a = as[k];
            (*br)[k].push_back(b);
        }
        return;
    }
    GroupArgs<R> ga;
    memset(&ga, 0, sizeof ga);
    ga.ncls = n;
    int w = 0;
    for (int k = 0; k < n; ++k) {
        const Geo &g = G.g[k];
        const int extra = (KIND == PK_KIN && (as[k].flags & PF_FINALIZE)) ? as[k].B.nb : 0;
        ga.blk0[k] = w;
        ga.nblk[k] = g.grid;
        ga.vpt[k] = g.vpt;
        ga.geo[k] = pass_geo(g.log2n, g.log2T, g.log2TW, as[k].P.pow2);
        ga.a[k] = as[k];
        w += (g.grid + extra + 7) / 8 * 8;  // the next class starts on a multiple of 8
    }
    ga.blk0[n] = w;
    descs.push_back(ga);
    out.push_back(GLaunch{(const void *)k_pass_group<R, KIND, DIR>, w, G.threads, G.lds, (int)descs.size() - 1});
}

// A context's device step counter (a replayed step sequence reads its field
// rows and records at the counter's step).
int64_t *member_step(qnd_ctx *c) { return reinterpret_cast<int64_t *>(c->d_small + SMALL_STEP); }

// After a group call: every member's slots in the axis-0 Fourier
// representation owing their mask, and the records copied out.
int group_finish(qnd_ctx *const *cs, int n, int64_t nsteps, bool want_obs, int width, double *const *obs_out)
{
    for (int k = 0; k < n; ++k) {
        qnd_ctx *c = cs[k];
        for (int b = 0; b < c->bnb; ++b) {
            c->slots[b].rep = REP_M;
            c->slots[b].pend_mask = nsteps > 0 ? (c->P.mask_on != 0) : c->slots[b].pend_mask;
        }
        if (want_obs && obs_out && obs_out[k])
            HIPCK(c, hipMemcpy(obs_out[k], c->d_rec, (size_t)c->bnb * nsteps * width * sizeof(double),
                               hipMemcpyDeviceToHost));
    }
    return 0;
}

// Steps of a group in the branch form, in chunks: step(first, k) enqueues k
// steps (the first of them as recorded for the call's first step when
// `first`).  The first chunk is that one step, then chunks of the steady
// step, eagerly or as captured HIP graphs of up to QND_GRAPH_STEPS steps each
// (one graph launch runs that many steps of every member).
#define QND_GRAPH_STEPS 32
template <class F> int branch_steps(qnd_ctx *c0, int64_t nsteps, bool use_graph, F &&step)
{
    int rc;
    int64_t done = 0;
    if (nsteps > 0) {
        if ((rc = step(true, 1))) return rc;
        done = 1;
    }
    if (use_graph && !c0->prof) {
        for (int64_t k = std::min<int64_t>(QND_GRAPH_STEPS, nsteps - done); k > 0;
             k = std::min<int64_t>(k, nsteps - done)) {
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            HIPCK(c0, hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal));
            const int erc = step(false, k);
            const hipError_t ec = hipStreamEndCapture(c0->stream, &graph);
            if (erc) {
                if (graph) hipGraphDestroy(graph);
                return erc;
            }
            if (ec != hipSuccess) return fail(c0, QND_EHIP, "graph capture failed: %s", hipGetErrorString(ec));
            hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (ei != hipSuccess) {
                hipGraphDestroy(graph);
                return fail(c0, QND_EHIP, "graph instantiation failed: %s", hipGetErrorString(ei));
            }
            hipError_t el = hipSuccess;
            for (; nsteps - done >= k && el == hipSuccess; done += k) el = hipGraphLaunch(exec, c0->stream);
            hipGraphExecDestroy(exec);
            hipGraphDestroy(graph);
            if (el != hipSuccess) return fail(c0, QND_EHIP, "graph launch failed: %s", hipGetErrorString(el));
        }
    }
    while (done < nsteps) {
        const int64_t k = std::min<int64_t>(QND_GRAPH_STEPS, nsteps - done);
        if ((rc = step(false, k))) return rc;
        done += k;
    }
    return 0;
}

template <typename R>
int group_real_steps(qnd_ctx *const *cs, int n, int64_t step0, int64_t nsteps, const int *codes, int ncodes,
                     double *const *obs_out, int use_graph)
{
    qnd_ctx *c0 = cs[0];
    const int d = c0->dim, s = c0->s;
    int rc;
    int width = 0;
    for (int k = 0; k < ncodes; ++k) width += obs_width(codes[k], d);
    const bool want_obs = width > 0;
    // each member: its batch in one state, in the axis-0 Fourier representation
    for (int k = 0; k < n; ++k) {
        qnd_ctx *c = cs[k];
        const int nb = c->bnb;
        bool same = true;
        for (int b = 1; b < nb; ++b) {
            const SlotState &a0 = c->slots[0], &ab = c->slots[b];
            if (ab.rep != a0.rep || ab.pend_mask != a0.pend_mask || ab.pend_vend != a0.pend_vend) same = false;
        }
        if (!same || c->slots[0].pend_vend)
            for (int b = 0; b < nb; ++b)
                if ((rc = ensure_rep<R>(c, b, REP_R))) return rc;
        if (c->slots[0].rep == REP_R) {
            if ((rc = plain_pass<R>(c, state_ptr<R>(c, 0), 0, -1, 1.0, nb))) return rc;
            for (int b = 0; b < nb; ++b) c->slots[b].rep = REP_M;
        }
        if (want_obs) {
            const size_t need = (size_t)nb * nsteps * width;
            if (need > c->rec_cap) {
                if (c->d_rec) HIPCK(c, hipFree(c->d_rec));
                c->d_rec = nullptr;
                c->rec_cap = 0;
                HIPCK(c, hipMalloc(&c->d_rec, need * sizeof(double)));
                c->rec_cap = need;
            }
        }
    }
    // the branch form (use_graph bit 1): each member's own kernels and
    // geometry on its own stream, the members' passes of a step concurrent
    const bool branch = (use_graph & 2) != 0;
    GroupGeo G0, G1;
    if (branch) {
        G0.g.resize(n);
        G1.g.resize(n);
        for (int k = 0; k < n; ++k) {
            G0.g[k] = geometry(cs[k], 0, cs[k]->bnb, PK_POT);
            G1.g[k] = geometry(cs[k], 1, cs[k]->bnb, PK_KIN);
        }
    } else if (!group_geometry(cs, n, 0, PK_POT, G0) || !group_geometry(cs, n, 1, PK_KIN, G1)) {
        return fail(c0, QND_EUNSUPPORTED, "no workgroup size fits every member of the group");
    }
    for (int k = 0; k < n; ++k) {
        qnd_ctx *c = cs[k];
        if ((rc = ensure_part(c, (size_t)G0.g[k].grid * QND_NOBS))) return rc;
        // tables the passes read, built before any capture
        if (G1.g[k].vpt == 4) {
            for (int ia = 0; ia <= s; ++ia) {
                const double ck = ia == 0 ? c->as_scaled[0]
                                          : c->as_scaled[2 * ia - 1] + (ia < s ? c->as_scaled[2 * ia] : 0.0);
                (void)ktab_for<R>(c, ck, G1.g[k], &rc);
                if (rc) return rc;
            }
        }
        if (G0.g[k].vpt == 4) {
            (void)vtab_for(c, &rc);
            if (rc) return rc;
        }
        HIPCK(c, hipStreamSynchronize(c->stream));  // member work on its own stream is done
    }
    // the device step counter (the first member's; in the branch form each
    // member advances its own)
    int64_t *d_step = member_step(c0);
    for (int k = 0; k < (branch ? n : 1); ++k) {
        HIPCK(c0, hipMemcpyAsync(member_step(cs[k]), &step0, sizeof(int64_t), hipMemcpyHostToDevice, cs[k]->stream));
        HIPCK(c0, hipStreamSynchronize(cs[k]->stream));
    }
    ObsSpec spec;
    memset(&spec, 0, sizeof spec);
    spec.ncodes = ncodes;
    for (int k = 0; k < ncodes; ++k) spec.codes[k] = codes[k];
    auto batch_rows = [&](PassArgs<R> &a, qnd_ctx *c) {
        a.B.nb = c->bnb;
        a.B.step = branch ? member_step(c) : d_step;
        a.B.step_base = step0;
    };
    // one step: boundary pass (owed mask, records), K(a_0) with the records'
    // finalize, then per stage the potential pass and the next kinetic pass
    // pend[k]: member k owes its mask (each member's own flag: the group
    // kernel reads flags per member).  record_step lists one step's launches:
    // boundary pass (owed mask, records), K(a_0) with the records' finalize,
    // then per stage the potential pass and the next kinetic pass, then the
    // step counter.
    std::vector<GroupArgs<R>> descs;
    auto record_step = [&](const std::vector<char> &pend, std::vector<GLaunch> &out, BLists<R> *br) -> int {
        std::vector<PassArgs<R>> as(n);
        bool boundary = want_obs;
        for (int k = 0; k < n; ++k) boundary |= pend[k] != 0;
        if (boundary) {
            for (int k = 0; k < n; ++k) {
                qnd_ctx *c = cs[k];
                PassArgs<R> &a = as[k];
                a = base_args<R>(c, state_ptr<R>(c, 0), G0.g[k]);
                batch_rows(a, c);
                a.scale1 = 1.0 / (double)c->nr[0];
                a.flags = PF_SECOND | (pend[k] ? PF_MASK : 0u);
                if (want_obs) {
                    a.flags |= PF_OBS;
                    a.partials = c->d_part;
                    a.npart = QND_NOBS;
                }
                if (c->round32) a.flags |= PF_ROUND32;
            }
            record_group<R, PK_POT, +1>(G0, as, descs, out, br);
        }
        auto kin = [&](int ia, bool fin) -> int {
            for (int k = 0; k < n; ++k) {
                qnd_ctx *c = cs[k];
                const double ck = ia == 0 ? c->as_scaled[0]
                                          : c->as_scaled[2 * ia - 1] + (ia < s ? c->as_scaled[2 * ia] : 0.0);
                PassArgs<R> &a = as[k];
                a = base_args<R>(c, state_ptr<R>(c, 0), G1.g[k]);
                batch_rows(a, c);
                a.kmode = KM_PHASE;
                a.kin_axes = 3u;
                a.ck_mid = ck;
                a.scale2 = 1.0 / (double)c->nr[1];
                int rr;
                a.ktab = ktab_for<R>(c, ck, G1.g[k], &rr);
                if (rr) return rr;
                if (a.ktab) a.flags |= PF_KTAB;
                else if (kax_pays(c, G1.g[k])) {
                    a.ktab = kax_for<R>(c, KM_PHASE, ck, &rr);
                    if (rr) return rr;
                    if (a.ktab) a.flags |= PF_KAX;
                }
                if (fin) {
                    a.flags |= PF_FINALIZE;
                    a.fin_part = c->d_part;
                    a.fin_nwg = G0.g[k].grid / c->bnb;
                    a.fin_rec = c->d_rec;
                    a.fin_spec = spec;
                    a.B.finE = c->d_bEo;
                    a.B.finE_step = d;
                    a.B.finE_stride = c->bntime * d;
                    a.B.rec_step = width;
                    a.B.rec_stride = nsteps * width;
                }
                if (c->round32) a.flags |= PF_ROUND32;
            }
            if (fin) record_group<R, PK_KIN, -1>(G1, as, descs, out, br);
            else record_group<R, PK_KINP, -1>(G1, as, descs, out, br);
            return 0;
        };
        int r;
        if ((r = kin(0, want_obs))) return r;
        for (int ia = 0; ia < s; ++ia) {
            for (int k = 0; k < n; ++k) {
                qnd_ctx *c = cs[k];
                PassArgs<R> &a = as[k];
                a = base_args<R>(c, state_ptr<R>(c, 0), G0.g[k]);
                batch_rows(a, c);
                a.scale1 = 1.0 / (double)c->nr[0];
                a.flags = PF_V1 | PF_SECOND;
                a.vmode = VM_LENGTH;
                a.cv1 = c->as_scaled[2 * ia] + c->as_scaled[2 * ia + 1];
                a.B.fld = c->d_bEt + (int64_t)ia * d;
                a.B.fld_step = (int64_t)s * d;
                a.B.fld_stride = c->bntime * s * d;
                if (G0.g[k].vpt == 4) {
                    int rr;
                    a.vtab = vtab_for(c, &rr);
                    if (rr) return rr;
                }
                if (c->round32) a.flags |= PF_ROUND32;
            }
            record_group<R, PK_POTL, +1>(G0, as, descs, out, br);
            if ((r = kin(ia + 1, false))) return r;
        }
        out.push_back(GLaunch{nullptr, 1, 64, 0, -1});
        return 0;
    };
    // the branch form: k steps of every member on its own stream, forked
    // from the first member's stream and joined back once; each member's
    // steps advance its own step counter, so the members proceed at their
    // own pace between the fork and the join
    hipEvent_t ev_fork = nullptr, ev_join[QND_GROUP_MAX] = {};
    BLists<R> bfirst(n), bsteady(n);  // the first and the steady step of each member
    auto run_branch = [&](bool first, int64_t k) -> int {
        HIPCK(c0, hipEventRecord(ev_fork, c0->stream));
        for (int m = 1; m < n; ++m) HIPCK(c0, hipStreamWaitEvent(cs[m]->stream, ev_fork, 0));
        for (int64_t i = 0; i < k; ++i) {
            const BLists<R> &Lb = (first && i == 0) ? bfirst : bsteady;
            size_t len = 0;
            for (int m = 0; m < n; ++m) len = std::max(len, Lb[m].size());
            for (size_t j = 0; j < len; ++j)  // the members' passes interleaved
                for (int m = 0; m < n; ++m) {
                    if (j >= Lb[m].size()) continue;
                    BLaunch<R> b = Lb[m][j];
                    void *args[] = {&b.a.data, &b.a.tw, &b.a.S, &b.geo, &b.fl, &b.nblk, &b.a};
                    HIPCK(c0, hipLaunchKernel(b.fn, dim3(b.grid), dim3(b.threads), args, b.lds, cs[m]->stream));
                }
            for (int m = 0; m < n; ++m)
                hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, cs[m]->stream, member_step(cs[m]));
            HIPCK(c0, hipGetLastError());
        }
        for (int m = 1; m < n; ++m) {
            HIPCK(c0, hipEventRecord(ev_join[m], cs[m]->stream));
            HIPCK(c0, hipStreamWaitEvent(c0->stream, ev_join[m], 0));
        }
        return 0;
    };
    const GroupArgs<R> *d_desc = nullptr;
    auto run_step = [&](const std::vector<GLaunch> &L) -> int {
        for (const GLaunch &l : L) {
            if (l.desc < 0) {
                hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, c0->stream, d_step);
            } else {
                const GroupArgs<R> *ptr = d_desc + l.desc;
                void *args[] = {(void *)&ptr};
                HIPCK(c0, hipLaunchKernel(l.fn, dim3(l.grid), dim3(l.threads), args, l.lds, c0->stream));
            }
            HIPCK(c0, hipGetLastError());
        }
        return 0;
    };
    std::vector<char> pend0(n), steady(n);  // owed masks now, and after every step
    for (int k = 0; k < n; ++k) {
        pend0[k] = cs[k]->slots[0].pend_mask ? 1 : 0;
        steady[k] = cs[k]->P.mask_on ? 1 : 0;
    }
    std::vector<GLaunch> first, steadyL;
    // the first step: what each member owes now
    if ((rc = record_step(pend0, first, branch ? &bfirst : nullptr))) return rc;
    if ((rc = record_step(steady, steadyL, branch ? &bsteady : nullptr))) return rc;
    if (branch) {
        for (int k = 1; k < n; ++k)
            if (cs[k]->stream == c0->stream) return fail(c0, QND_EINVAL, "group members share a stream");
        struct Events {  // the fork and join events, destroyed on every way out
            hipEvent_t *fork, *join;
            int n;
            ~Events()
            {
                if (*fork) (void)hipEventDestroy(*fork);
                for (int k = 1; k < n; ++k)
                    if (join[k]) (void)hipEventDestroy(join[k]);
            }
        } ev_guard{&ev_fork, ev_join, n};
        HIPCK(c0, hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        for (int k = 1; k < n; ++k) HIPCK(c0, hipEventCreateWithFlags(&ev_join[k], hipEventDisableTiming));
        rc = branch_steps(c0, nsteps, (use_graph & 1) != 0,
                          [&](bool first, int64_t k) { return run_branch(first, k); });
        if (rc) return rc;
        HIPCK(c0, hipStreamSynchronize(c0->stream));
        return group_finish(cs, n, nsteps, want_obs, width, obs_out);
    }
    {   // the call's descriptors, uploaded once (they serve every step)
        const size_t bytes = descs.size() * sizeof(GroupArgs<R>);
        if (bytes > c0->gdesc_cap) {
            if (c0->d_gdesc) HIPCK(c0, hipFree(c0->d_gdesc));
            c0->d_gdesc = nullptr;
            c0->gdesc_cap = 0;
            HIPCK(c0, hipMalloc(&c0->d_gdesc, bytes));
            c0->gdesc_cap = bytes;
        }
        HIPCK(c0, hipMemcpy(c0->d_gdesc, descs.data(), bytes, hipMemcpyHostToDevice));
        d_desc = reinterpret_cast<const GroupArgs<R> *>(c0->d_gdesc);
    }
    int64_t done = 0;
    if (nsteps > 0) {
        if ((rc = run_step(first))) return rc;
        done = 1;
    }
    bool prof = false;
    for (int k = 0; k < n; ++k) prof |= cs[k]->prof;
    if ((use_graph & 1) && nsteps - done > 0 && !prof) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        HIPCK(c0, hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal));
        const int erc = run_step(steadyL);
        const hipError_t ec = hipStreamEndCapture(c0->stream, &graph);
        if (erc) {
            if (graph) hipGraphDestroy(graph);
            return erc;
        }
        if (ec != hipSuccess) return fail(c0, QND_EHIP, "graph capture failed: %s", hipGetErrorString(ec));
        hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            hipGraphDestroy(graph);
            return fail(c0, QND_EHIP, "graph instantiation failed: %s", hipGetErrorString(ei));
        }
        hipError_t el = hipSuccess;
        for (; done < nsteps && el == hipSuccess; ++done) el = hipGraphLaunch(exec, c0->stream);
        hipGraphExecDestroy(exec);
        hipGraphDestroy(graph);
        if (el != hipSuccess) return fail(c0, QND_EHIP, "graph launch failed: %s", hipGetErrorString(el));
    }
    for (; done < nsteps; ++done)
        if ((rc = run_step(steadyL))) return rc;
    HIPCK(c0, hipStreamSynchronize(c0->stream));
    return group_finish(cs, n, nsteps, want_obs, width, obs_out);
}

// Write_final_observables numerics (src/io_functions.c:159-199)
template <typename R>
int final_observables(qnd_ctx *c, int slot, const double *R_mask, const double *pr_max, int Ne, double de,
                      double *ped, double *psi_p)
{
    int rc = ensure_rep<R>(c, slot, REP_R);
    if (rc) return rc;
    C<R> *psi = state_ptr<R>(c, slot);
    C<R> *A = reinterpret_cast<C<R> *>(c->d_scr[0]);
    const int64_t N = c->N;
    // radial sin^2 mask into a scratch copy (the slot itself is unchanged)
    double *r1 = c->d_small + SMALL_SUMS + 128;
    HIPCK(c, hipMemcpyAsync(r1, &R_mask[1], sizeof(double), hipMemcpyHostToDevice, c->stream));
    {
        ProfScope ps(c, PC_SMALL);
        hipLaunchKernelGGL(k_elem<R>, dim3(flat_grid(c->N, 4, 1024, 16384)), dim3(256),
                           0, c->stream, (int)EM_RMASK, A, (const C<R> *)psi, 0,
                           R_mask[0], N, c->P, (const C<R> *)nullptr, (int64_t)0, 0, (const double *)r1);
        HIPCK(c, hipGetLastError());
    }
    std::vector<C<R>> h(psi_p ? N : 0);
    if ((rc = full_fft<R>(c, A, -1))) return rc;
    double fs = 1.0;
    for (int j = 0; j < c->dim; ++j) fs /= (double)c->nr[j];
    if ((rc = elem<R>(c, EM_SCALE, A, A, 0, std::sqrt(fs)))) return rc;
    if (ped && Ne > 0) {
        const int chunk = 1024;
        const int grid = (int)((N + chunk - 1) / chunk);
        if ((rc = ensure_part(c, (size_t)grid * Ne))) return rc;
        double pm[QND_MAXD] = {0, 0, 0, 0};
        for (int j = 0; j < c->dim; ++j) pm[j] = pr_max[j];
        if (!c->d_prmax) HIPCK(c, hipMalloc(&c->d_prmax, QND_MAXD * sizeof(double)));
        HIPCK(c, hipMemcpyAsync(c->d_prmax, pm, sizeof pm, hipMemcpyHostToDevice, c->stream));
        const double sigma = de;
        const double nf = 1. / std::pow(2. * M_PI * sigma * sigma, .5 * c->dim);
        hipLaunchKernelGGL(k_ped<R>, dim3(grid), dim3(256), 0, c->stream, (const C<R> *)A, N, c->P,
                           (const double *)c->d_prmax, Ne, de, sigma, nf, chunk, c->d_part);
        HIPCK(c, hipGetLastError());
        double *sums = nullptr;
        HIPCK(c, hipMalloc(&sums, Ne * sizeof(double)));
        hipLaunchKernelGGL(k_sum_partials, dim3((Ne + 3) / 4), dim3(256), 0, c->stream, c->d_part, grid, Ne, Ne, 1.0,
                           sums);
        HIPCK(c, hipGetLastError());
        HIPCK(c, hipMemcpyAsync(ped, sums, Ne * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCK(c, hipStreamSynchronize(c->stream));
        HIPCK(c, hipFree(sums));
    }
    if (psi_p) {
        HIPCK(c, hipMemcpyAsync(h.data(), A, N * sizeof(C<R>), hipMemcpyDeviceToHost, c->stream));
        HIPCK(c, hipStreamSynchronize(c->stream));
        for (int64_t f = 0; f < N; ++f) {
            psi_p[2 * f] = (double)h[f].x;
            psi_p[2 * f + 1] = (double)h[f].y;
        }
    }
    return 0;
}

template <typename R> int upload(qnd_ctx *c, int slot, const double *psi)
{
    const int64_t N = c->N;
    C<R> *dst = state_ptr<R>(c, slot);
    if (sizeof(R) == 8) {
        HIPCK(c, hipMemcpyAsync(dst, psi, N * 16, hipMemcpyHostToDevice, c->stream));
    } else {
        std::vector<C<R>> h(N);
        for (int64_t f = 0; f < N; ++f) {
            h[f].x = (R)psi[2 * f];
            h[f].y = (R)psi[2 * f + 1];
        }
        HIPCK(c, hipMemcpyAsync(dst, h.data(), N * sizeof(C<R>), hipMemcpyHostToDevice, c->stream));
    }
    HIPCK(c, hipStreamSynchronize(c->stream));
    c->slots[slot] = SlotState();
    return 0;
}

template <typename R> int download(qnd_ctx *c, int slot, double *psi)
{
    int rc = ensure_rep<R>(c, slot, REP_R);
    if (rc) return rc;
    const int64_t N = c->N;
    C<R> *src = state_ptr<R>(c, slot);
    if (sizeof(R) == 8) {
        HIPCK(c, hipMemcpyAsync(psi, src, N * 16, hipMemcpyDeviceToHost, c->stream));
        HIPCK(c, hipStreamSynchronize(c->stream));
    } else {
        std::vector<C<R>> h(N);
        HIPCK(c, hipMemcpyAsync(h.data(), src, N * sizeof(C<R>), hipMemcpyDeviceToHost, c->stream));
        HIPCK(c, hipStreamSynchronize(c->stream));
        for (int64_t f = 0; f < N; ++f) {
            psi[2 * f] = (double)h[f].x;
            psi[2 * f + 1] = (double)h[f].y;
        }
    }
    return 0;
}

template <typename R> int make_twiddles(qnd_ctx *c, int axis)
{
    AxisPlan &p = c->plan[axis];
    for (int j = 0; j < axis; ++j)
        if (c->plan[j].n == p.n && c->plan[j].tw) {  // share
            p.tw = nullptr;
            p.tw = c->plan[j].tw;
            return 0;
        }
    // W_n^k = exp(-2 pi i k/n) from long double; complex64 keeps the rounding
    // remainder of each component as a second float (hi + lo)
    using TT = typename TwT<R>::T;
    std::vector<TT> t(p.n);
    for (int k = 0; k < p.n; ++k) {
        const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)p.n;
        const long double cs = cosl(ang), sn = sinl(ang);
        if constexpr (sizeof(R) == 8) {
            t[k].x = (double)cs;
            t[k].y = (double)sn;
        } else {
            const float ch = (float)cs, sh = (float)sn;
            t[k].x = ch;
            t[k].y = sh;
            t[k].z = (float)(cs - (long double)ch);
            t[k].w = (float)(sn - (long double)sh);
        }
    }
    HIPCK(c, hipMalloc(&p.tw, p.n * sizeof(TT)));
    HIPCK(c, hipMemcpy(p.tw, t.data(), p.n * sizeof(TT), hipMemcpyHostToDevice));
    return 0;
}

// qnd_observables_noalloc: the fused position-space observables (norm,
// dipole_r, dipole_a) of a caller's device state with the reduction of
// qnd_observables (same kernel, same partition into workgroups, same column
// sums), its partials and sums in this device's scratch array, and the record
// assembled on the device into the caller's output.  Nothing is allocated.
__global__ void k_obs_assemble(const double *__restrict__ sums, ObsSpec spec, int dim, double *out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int off = 0;
    for (int k = 0; k < spec.ncodes; ++k) {
        const int c = spec.codes[k];
        if (c == QND_OBS_NORM) out[off++] = sums[0];
        else if (c == QND_OBS_DIPOLE_R)
            for (int j = 0; j < dim; ++j) out[off++] = sums[1 + j];
        else if (c == QND_OBS_DIPOLE_A)
            for (int j = 0; j < dim; ++j) out[off++] = -sums[1 + QND_MAXD + j] - spec.E[j] * sums[0];
    }
}

template <typename R>
int observables_noalloc(qnd_ctx *c, const void *psi_dev, const ObsSpec &spec, double *out_dev)
{
    double *scr = (c->device >= 0 && c->device < QND_MAX_DEVICES) ? g_noalloc_ptr[c->device] : nullptr;
    if (!scr)
        return fail(c, QND_EHIP, "device scratch of the allocation-free entry point unavailable on device %d (%s)",
                    c->device,
                    c->device >= 0 && c->device < QND_MAX_DEVICES ? hipGetErrorString(g_noalloc_err[c->device])
                                                                  : "device index out of range");
    const int grid = flat_grid(c->N, 8, 512, 8192), npart = QND_NOBS;
    if ((size_t)grid * npart + QND_NOBS > (size_t)QND_NOALLOC_DOUBLES)
        return fail(c, QND_EINVAL, "%d partial rows exceed the %d rows of the allocation-free scratch", grid,
                    QND_NOALLOC_ROWS);
    double *part = scr, *sums = scr + (size_t)QND_NOALLOC_ROWS * QND_NOBS;
    hipLaunchKernelGGL(k_red<R>, dim3(grid), dim3(256), 0, c->stream, (int)RM_OBS, (const C<R> *)psi_dev,
                       (C<R> *)nullptr, (const C<R> *)nullptr, 0, 0.0, c->N, c->P, (const double *)nullptr, part,
                       npart);
    HIPCK(c, hipGetLastError());
    hipLaunchKernelGGL(k_sum_partials, dim3((QND_NOBS + 3) / 4), dim3(256), 0, c->stream, (const double *)part,
                       grid, npart, QND_NOBS, c->P.dNr, sums);
    HIPCK(c, hipGetLastError());
    hipLaunchKernelGGL(k_obs_assemble, dim3(1), dim3(64), 0, c->stream, (const double *)sums, spec, c->dim,
                       out_dev);
    HIPCK(c, hipGetLastError());
    return 0;
}

bool potential_params(const char *name, int *kind, double *b, double *c2)
{
    *kind = 0;
    *b = 0.0;
    if (!strcmp(name, "H_1D")) *c2 = 2.;
    else if (!strcmp(name, "H_3D")) *c2 = 1e-4;
    else if (!strcmp(name, "H_2D")) *c2 = .7396;
    else if (!strcmp(name, "Ne_2D")) { *b = 9.; *c2 = 2.88172; }
    else if (!strcmp(name, "Ar_2D")) { *b = 17.; *c2 = .91205; }
    else if (!strcmp(name, "Kr_2D")) { *b = 35.; *c2 = .7875; }
    else if (!strcmp(name, "Xe_2D")) { *b = 53.; *c2 = .52592; }
    else if (!strcmp(name, "Ne_short_2D")) { *kind = 1; *c2 = 0.; }
    else return false;
    return true;
}

int check_ready(qnd_ctx *c)
{
    if (!c) return QND_EINVAL;
    if (!c->have_pot) return fail(c, QND_ESTATE, "potential not set (qnd_set_potential)");
    if (!c->have_int) return fail(c, QND_ESTATE, "integrator not set (qnd_set_integrator)");
    return 0;
}

int check_slot(qnd_ctx *c, int slot)
{
    if (slot < 0 || slot >= c->n_slots) return fail(c, QND_EINVAL, "slot %d out of range [0,%d)", slot, c->n_slots);
    return 0;
}

}  // namespace

// =====================================================================
// C ABI
// =====================================================================
#define DISPATCH(c, fn, ...) ((c)->prec == QND_C64 ? fn<float>(__VA_ARGS__) : fn<double>(__VA_ARGS__))

extern "C" {

const char *qnd_version(void) { return QND_VERSION; }

const char *qnd_last_error(const qnd_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int qnd_create(qnd_ctx **out, int dim, const int64_t *nr, const double *dr, int precision, int n_slots)
{
    if (!out) return fail(nullptr, QND_EINVAL, "null output pointer");
    *out = nullptr;
    if (dim < 1 || dim > QND_MAXD) return fail(nullptr, QND_EINVAL, "dimension %d not in [1,%d]", dim, QND_MAXD);
    if (precision != QND_C128 && precision != QND_C64 && precision != QND_C64F64)
        return fail(nullptr, QND_EINVAL, "unknown precision %d", precision);
    if (n_slots < 1) return fail(nullptr, QND_EINVAL, "n_slots must be >= 1");
    for (int j = 0; j < dim; ++j) {
        const int64_t m = (nr[j] % 3 == 0) ? nr[j] / 3 : nr[j];
        const bool ok = (nr[j] % 3 == 0) ? (m >= 4 && (m & (m - 1)) == 0) : (m >= 8 && (m & (m - 1)) == 0);
        if (!ok)
            return fail(nullptr, QND_EUNSUPPORTED,
                        "NR[%d] = %lld: the GPU path needs a power of two >= 8 or 3 * 2^k >= 12", j,
                        (long long)nr[j]);
        if (!(dr[j] > 0)) return fail(nullptr, QND_EINVAL, "DR[%d] must be positive", j);
    }
    qnd_ctx *c = new qnd_ctx();
    c->dim = dim;
    // complex64 storage with FP64 pass arithmetic is a complex64 context
    // (tables, buffers, host paths) whose passes are the k_pass_x kernels
    c->x64 = precision == QND_C64F64;
    c->prec = c->x64 ? QND_C64 : precision;
    c->esz = c->prec == QND_C64 ? 8 : 16;
    c->lds_esz = c->x64 ? 16 : c->esz;
    c->n_slots = n_slots;
    c->N = 1;
    for (int j = 0; j < dim; ++j) {
        c->nr[j] = nr[j];
        c->dr[j] = dr[j];
        c->N *= nr[j];
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        delete c;
        return fail(nullptr, QND_EHIP, "no HIP device available");
    }
    hipGetDevice(&c->device);
    noalloc_setup(c->device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) {
        // dynamic LDS per workgroup, less room for the kernels' static arrays
        size_t m = prop.sharedMemPerBlock;
        if (m > 163840) m = 163840;
        c->lds_max = m > 4096 ? m - 4096 : m;
        if (prop.multiProcessorCount > 0) c->n_cu = prop.multiProcessorCount;
    }
    for (int j = 0; j < dim; ++j) {
        const int64_t tpl = (nr[j] % 3 == 0) ? nr[j] / 12 : nr[j] / 8;  // threads per line
        if (2 * (size_t)nr[j] * c->lds_esz > c->lds_max || tpl > QND_PASS_THREADS) {
            delete c;
            return fail(nullptr, QND_EUNSUPPORTED, "NR[%d] = %lld exceeds the single-pass line limit", j,
                        (long long)nr[j]);
        }
    }
    // physics defaults
    Phys &P = c->P;
    memset(&P, 0, sizeof P);
    P.dim = dim;
    P.dNr = 1.0;
    P.pow2 = 1;
    for (int j = 0; j < QND_MAXD; ++j) P.n[j] = 1;
    int off = 0;
    for (int j = 0; j < dim; ++j) {
        P.n[j] = (int)nr[j];
        P.lg[j] = ilog2(nr[j]);
        if ((nr[j] & (nr[j] - 1)) != 0) P.pow2 = 0;
        P.off[j] = off;
        off += (int)nr[j];
        P.dr[j] = dr[j];
        P.L[j] = dr[j] * nr[j];
        P.dp[j] = 2. * M_PI / P.L[j];
        P.dNr *= dr[j];
    }
    if (const char *e = getenv("QND_TW")) {  // tuning knob: "contiguous,strided" lines per workgroup
        int a = 0, b = 0;
        if (sscanf(e, "%d,%d", &a, &b) >= 1) {
            c->tw_override[0] = a;
            c->tw_override[1] = b;
        }
    }
    if (const char *e = getenv("QND_WAVE_LINES")) c->wave_lines = atoi(e) != 0;
    if (const char *e = getenv("QND_ROUND_STATE")) c->round32 = precision == QND_C128 && atoi(e) != 0;
    if (const char *e = getenv("QND_KAX")) c->use_kax = atoi(e) != 0;
    if (const char *e = getenv("QND_IMAG1D")) c->imag1d = atoi(e) != 0;
    if (const char *e = getenv("QND_GS_SPLIT")) c->gs_split = atoi(e) != 0 ? 1 : 0;
    if (const char *e = getenv("QND_VPT")) {  // tuning knob: values per thread
        const int v = atoi(e);
        if (v == 4 || v == 8) c->vpt_override = v;
    }
    c->slots.resize(n_slots);
    int rc = 0;
    auto bail = [&](int code) {
        qnd_destroy(c);
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(nullptr, QND_EHIP, "hipStreamCreate failed"));
    size_t sbytes = (size_t)c->N * c->esz;
    if (hipMalloc(&c->d_states, sbytes * n_slots) != hipSuccess)
        return bail(fail(nullptr, QND_EHIP, "hipMalloc of %zu bytes for states failed", sbytes * n_slots));
    hipMemset(c->d_states, 0, sbytes * n_slots);
    for (int k = 0; k < 3; ++k)
        if (hipMalloc(&c->d_scr[k], sbytes) != hipSuccess)
            return bail(fail(nullptr, QND_EHIP, "hipMalloc of scratch failed"));
    if (hipMalloc(&c->d_small, SMALL_TOTAL * sizeof(double)) != hipSuccess)
        return bail(fail(nullptr, QND_EHIP, "hipMalloc failed"));
    hipMemset(c->d_small, 0, SMALL_TOTAL * sizeof(double));
    if (hipMalloc(&c->d_capt, off * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->d_maskt, off * sizeof(double)) != hipSuccess)
        return bail(fail(nullptr, QND_EHIP, "hipMalloc failed"));
    hipMemset(c->d_capt, 0, off * sizeof(double));   // read unconditionally by the passes
    hipMemset(c->d_maskt, 0, off * sizeof(double));
    P.capt = c->d_capt;
    P.maskt = c->d_maskt;
    for (int j = 0; j < dim; ++j) {
        c->plan[j].n = (int)nr[j];
        c->plan[j].log2n = ilog2(nr[j]);
        c->plan[j].T = (int)nr[j] / 8;
        rc = c->prec == QND_C64 ? make_twiddles<float>(c, j) : make_twiddles<double>(c, j);
        if (rc) {
            g_create_err = c->err;
            return bail(rc);
        }
    }
    for (int k = 0; k < 16; ++k) hipEventCreate(&c->ev[k]);
    if (hipHostMalloc((void **)&c->h_gs, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_gs[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_gs[1], hipEventDisableTiming) != hipSuccess)
        return bail(fail(nullptr, QND_EHIP, "pinned host buffer or events for the Gram-Schmidt check failed"));
    *out = c;
    return 0;
}

void qnd_destroy(qnd_ctx *c)
{
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    prof_flush(c);
    for (int j = 0; j < QND_MAXD; ++j) {
        if (!c->plan[j].tw) continue;
        bool shared = false;
        for (int k = 0; k < j; ++k)
            if (c->plan[k].tw == c->plan[j].tw) shared = true;
        if (!shared) hipFree(c->plan[j].tw);
    }
    if (c->d_states) hipFree(c->d_states);
    for (int k = 0; k < 3; ++k)
        if (c->d_scr[k]) hipFree(c->d_scr[k]);
    if (c->d_part) hipFree(c->d_part);
    if (c->d_small) hipFree(c->d_small);
    if (c->d_rec) hipFree(c->d_rec);
    if (c->d_prmax) hipFree(c->d_prmax);
    if (c->d_capt) hipFree(c->d_capt);
    if (c->d_maskt) hipFree(c->d_maskt);
    if (c->d_bEt) hipFree(c->d_bEt);
    if (c->d_bAt) hipFree(c->d_bAt);
    if (c->d_bEo) hipFree(c->d_bEo);
    if (c->d_ktab) hipFree(c->d_ktab);
    if (c->d_atab) hipFree(c->d_atab);
    if (c->d_kax) hipFree(c->d_kax);
    if (c->d_Eo) hipFree(c->d_Eo);
    if (c->d_side) hipFree(c->d_side);
    if (c->d_kcoef) hipFree(c->d_kcoef);
    if (c->d_vtab) hipFree(c->d_vtab);
    if (c->d_gdesc) hipFree(c->d_gdesc);
    for (void *p : {c->d_adj_ckpt, c->d_adj_final, c->d_adj_save, (void *)c->d_adj_part})
        if (p) hipFree(p);
    for (int k = 0; k < 16; ++k)
        if (c->ev[k]) hipEventDestroy(c->ev[k]);
    for (int k = 0; k < 2; ++k)
        if (c->ev_gs[k]) hipEventDestroy(c->ev_gs[k]);
    if (c->h_gs) hipHostFree(c->h_gs);
    if (c->stream2) {
        hipStreamSynchronize(c->stream2);
        hipStreamDestroy(c->stream2);
    }
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int qnd_set_potential(qnd_ctx *c, const char *name, double softcore, const double *cap)
{
    if (!c || !name) return QND_EINVAL;
    if (int rc = flush_owed(c)) return rc;  // owed operators keep the set-up they were owed under
    ++c->setup_gen;
    int kind;
    double b, c2;
    if (!potential_params(name, &kind, &b, &c2))
        return fail(c, QND_EUNKNOWN_NAME, "Unknown potential: %s", name);
    if (softcore >= 0) c2 = softcore * softcore;
    c->P.pot_kind = kind;
    c->P.pot_b = b;
    c->P.pot_c2 = c2;
    ++c->pot_gen;
    // separable CAP: Vcap = sum_j cap0 (|x_j| - L_j)^2 for |x_j| >= L_j,
    // L_j = n_j dr_j / 2 - width (src/meshgrid.c:132-136)
    const double cap0 = cap ? cap[0] : 0.0, cap1 = cap ? cap[1] : 0.0;
    std::vector<double> t(c->P.off[c->dim - 1] + c->nr[c->dim - 1], 0.0);
    for (int j = 0; j < c->dim; ++j) {
        const double Lc = .5 * c->nr[j] * c->dr[j] - cap1;
        for (int64_t i = 0; i < c->nr[j]; ++i) {
            const double x = -.5 * (c->dr[j] * c->nr[j]) + i * c->dr[j];
            if (fabs(x) >= Lc) t[c->P.off[j] + i] = cap0 * (fabs(x) - Lc) * (fabs(x) - Lc);
        }
    }
    HIPCK(c, hipMemcpy(c->d_capt, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    c->P.cap_on = cap0 != 0.0 ? 1 : 0;
    c->have_pot = true;
    return 0;
}

int qnd_set_integrator(qnd_ctx *c, const char *name, double dt)
{
    if (!c || !name) return QND_EINVAL;
    if (int rc = flush_owed(c)) return rc;  // owed operators keep the set-up they were owed under
    ++c->setup_gen;
    static const double bm4[6] = {.0792036964311957, .1303114101821663, .2228614958676077,
                                  -.3667132690474257, .3246481886897062, .1096884778767498};
    static const double bm6[10] = {.050262764400392, .098553683500650, .314960616927694, -.447346482695478,
                                   .492426372489876, -.425118767797691, .237063913978122, .195602488600053,
                                   .346358189850727, -.362762779254345};
    static const double rkn4[6] = {.082984406417405, .162314550766866, .233995250731502,
                                   .370877414979578, -.409933719901926, .059762097006575};
    double as[20] = {0};
    int s;
    if (!strcmp(name, "Verlet")) { s = 1; as[0] = .5; }
    else if (!strcmp(name, "BM4")) { s = 6; memcpy(as, bm4, sizeof bm4); }
    else if (!strcmp(name, "BM6")) { s = 10; memcpy(as, bm6, sizeof bm6); }
    else if (!strcmp(name, "RKN4")) { s = 6; memcpy(as, rkn4, sizeof rkn4); }
    else return fail(c, QND_EUNKNOWN_NAME, "Unknown integrator: %s", name);
    for (int j = 0; j < s; ++j) as[s + j] = as[s - 1 - j];  // mirrored second half
    c->s = s;
    c->dt = dt;
    for (int j = 0; j < 20; ++j) {
        c->as[j] = as[j];
        c->as_scaled[j] = as[j] * dt;
    }
    c->have_int = true;
    return 0;
}

int qnd_set_mask(qnd_ctx *c, double frac)
{
    if (!c) return QND_EINVAL;
    if (int rc = flush_owed(c)) return rc;  // owed operators keep the set-up they were owed under
    ++c->setup_gen;
    // m_j(x) = cos(pi/2 (|x| - x0)/(L - x0))^(1/8) for |x| > x0 = (1-frac) L,
    // L = n_j dr_j / 2; the mask is the product over axes
    c->mask_frac = frac;
    std::vector<double> t(c->P.off[c->dim - 1] + c->nr[c->dim - 1], 1.0);
    for (int j = 0; j < c->dim && frac > 0; ++j) {
        const double L = .5 * c->nr[j] * c->dr[j], x0 = (1. - frac) * L;
        for (int64_t i = 0; i < c->nr[j]; ++i) {
            const double ax = fabs(-.5 * (c->dr[j] * c->nr[j]) + i * c->dr[j]);
            if (ax > x0) t[c->P.off[j] + i] = pow(cos(.5 * M_PI * (ax - x0) / (L - x0)), .125);
        }
    }
    HIPCK(c, hipMemcpy(c->d_maskt, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    c->P.mask_on = frac > 0 ? 1 : 0;
    return 0;
}

int qnd_upload_state(qnd_ctx *c, int slot, const double *psi)
{
    if (!c || !psi) return QND_EINVAL;
    int rc = check_slot(c, slot);
    if (rc) return rc;
    return DISPATCH(c, upload, c, slot, psi);
}

int qnd_download_state(qnd_ctx *c, int slot, double *psi)
{
    if (!c || !psi) return QND_EINVAL;
    int rc = check_slot(c, slot);
    if (rc) return rc;
    if (c->slots[slot].pend_vend && !c->have_int) return fail(c, QND_ESTATE, "integrator not set");
    return DISPATCH(c, download, c, slot, psi);
}

int qnd_transform(qnd_ctx *c, int slot, int sign)
{
    if (!c) return QND_EINVAL;
    int rc = check_slot(c, slot);
    if (rc) return rc;
    if (sign != -1 && sign != 1) return fail(c, QND_EINVAL, "sign must be -1 or +1");
    if (c->slots[slot].pend_vend && !c->have_int) return fail(c, QND_ESTATE, "integrator not set");
    rc = c->prec == QND_C64 ? ensure_rep<float>(c, slot, REP_R) : ensure_rep<double>(c, slot, REP_R);
    if (rc) return rc;
    rc = c->prec == QND_C64 ? full_fft<float>(c, state_ptr<float>(c, slot), sign)
                            : full_fft<double>(c, state_ptr<double>(c, slot), sign);
    if (rc) return rc;
    c->slots[slot] = SlotState();  // the content is now whatever the caller asked for
    HIPCK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int qnd_imag_steps(qnd_ctx *c, int slot, int64_t nsteps, int gram_schmidt)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_slot(c, slot))) return rc;
    if (nsteps < 0) return fail(c, QND_EINVAL, "nsteps < 0");
    if (gram_schmidt && slot > 64) return fail(c, QND_EUNSUPPORTED, "at most 65 states for Gram-Schmidt");
    return DISPATCH(c, imag_steps, c, slot, nsteps, gram_schmidt);
}

int qnd_gram_schmidt(qnd_ctx *c, int slot)
{
    if (!c) return QND_EINVAL;
    int rc = check_slot(c, slot);
    if (rc) return rc;
    return DISPATCH(c, gram_schmidt, c, slot);
}

int qnd_energy_moments(qnd_ctx *c, int slot, int M, double *E)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_slot(c, slot))) return rc;
    if (M <= 0) return 0;
    rc = DISPATCH(c, energies, c, slot, M, E);
    prof_flush(c);
    return rc;
}

int qnd_observables(qnd_ctx *c, int slot, const int *codes, int ncodes, const double *alpha,
                    const double *Efield, double *out)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_slot(c, slot))) return rc;
    rc = DISPATCH(c, observables, c, slot, codes, ncodes, alpha, Efield, out);
    prof_flush(c);
    return rc;
}

int qnd_observables_noalloc(qnd_ctx *c, const void *psi_dev, const int *codes, int ncodes, const double *Efield,
                            double *out_dev)
{
    if (!c) return QND_EINVAL;
    if (!c->have_pot) return fail(c, QND_ESTATE, "potential not set (qnd_set_potential)");
    if (!psi_dev || !out_dev || (ncodes > 0 && !codes)) return fail(c, QND_EINVAL, "null state, output or codes");
    if (ncodes < 0 || ncodes > 8) return fail(c, QND_EINVAL, "between 0 and 8 observables per call");
    ObsSpec spec;
    memset(&spec, 0, sizeof spec);
    spec.ncodes = ncodes;
    for (int k = 0; k < ncodes; ++k) {
        if (codes[k] != QND_OBS_NORM && codes[k] != QND_OBS_DIPOLE_R && codes[k] != QND_OBS_DIPOLE_A)
            return fail(c, QND_EUNSUPPORTED, "observable code %d: the allocation-free entry point records norm, "
                        "dipole_r and dipole_a only", codes[k]);
        spec.codes[k] = codes[k];
    }
    for (int j = 0; j < c->dim; ++j) spec.E[j] = Efield ? Efield[j] : 0.0;
    if (ncodes == 0) return 0;
    return DISPATCH(c, observables_noalloc, c, psi_dev, spec, out_dev);
}

int qnd_set_field(qnd_ctx *c, int64_t ntime, const double *Et, const double *At, const double *Eobs,
                  const double *Aobs)
{
    if (!c || ntime < 0) return QND_EINVAL;
    if (!c->have_int) return fail(c, QND_ESTATE, "integrator not set (qnd_set_integrator)");
    const size_t ns = (size_t)ntime * c->s * c->dim, no = (size_t)ntime * c->dim;
    c->Et.assign(ns, 0.0);
    c->At.assign(ns, 0.0);
    c->Eobs.assign(no, 0.0);
    c->Aobs.assign(no, 0.0);
    if (Et) memcpy(c->Et.data(), Et, ns * sizeof(double));
    if (At) memcpy(c->At.data(), At, ns * sizeof(double));
    if (Eobs) memcpy(c->Eobs.data(), Eobs, no * sizeof(double));
    if (Aobs) memcpy(c->Aobs.data(), Aobs, no * sizeof(double));
    // E(t_n) on the device: the field of the records that observable passes
    // over saved states finalize
    if (c->d_Eo) HIPCK(c, hipFree(c->d_Eo));
    c->d_Eo = nullptr;
    HIPCK(c, hipMalloc(&c->d_Eo, (no ? no : 1) * sizeof(double)));
    if (no) HIPCK(c, hipMemcpy(c->d_Eo, c->Eobs.data(), no * sizeof(double), hipMemcpyHostToDevice));
    c->ntime = ntime;
    c->field_s = c->s;
    ++c->setup_gen;
    return 0;
}

int qnd_real_steps(qnd_ctx *c, int slot, int gauge, int64_t step0, int64_t nsteps, const int *codes,
                   int ncodes, double *obs_out)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_slot(c, slot))) return rc;
    if (gauge != QND_GAUGE_LENGTH && gauge != QND_GAUGE_VELOCITY)
        return fail(c, QND_EINVAL, "unknown gauge %d", gauge);
    if (ncodes > 0 && (!codes || !obs_out)) return fail(c, QND_EINVAL, "observable codes need an output array");
    return DISPATCH(c, real_steps, c, slot, gauge, step0, nsteps, codes, ncodes, obs_out);
}

int qnd_forward_checkpointed(qnd_ctx *c, int slot, int64_t nsteps, int64_t every, double *obs_out,
                             double *norm_final)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_slot(c, slot))) return rc;
    if (c->dim < 2) return fail(c, QND_EUNSUPPORTED, "the adjoint needs dim >= 2");
    if (nsteps < 1 || every < 1) return fail(c, QND_EINVAL, "nsteps and the checkpoint interval must be >= 1");
    if (nsteps > c->ntime) return fail(c, QND_EINVAL, "%lld steps exceed the field table", (long long)nsteps);
    if (c->field_s != c->s)
        return fail(c, QND_ESTATE, "the field tables were laid out for %d stages but the integrator has %d "
                    "(set the field again after qnd_set_integrator)", c->field_s, c->s);
    return DISPATCH(c, adj_forward, c, slot, nsteps, every, obs_out, norm_final);
}

int qnd_adjoint(qnd_ctx *c, int64_t nsteps, const double *dL_da, double dL_dnorm, double *dL_dEt,
                double *dL_dEobs, double *dL_dsoft)
{
    if (!c || !dL_da || !dL_dEt || !dL_dEobs) return QND_EINVAL;
    if (c->adj_nsteps < 1 || c->adj_prec != c->prec)
        return fail(c, QND_ESTATE, "no checkpointed forward run (qnd_forward_checkpointed)");
    // the caller's arrays are sized for nsteps rows: they must be the
    // forward run's, or the sweep would read and write past them
    if (nsteps != c->adj_nsteps)
        return fail(c, QND_EINVAL, "the gradient arrays hold %lld steps but the checkpointed forward run has %lld",
                    (long long)nsteps, (long long)c->adj_nsteps);
    if (c->adj_gen != c->setup_gen)
        return fail(c, QND_ESTATE, "potential, integrator, mask or field changed since the checkpointed forward run");
    return DISPATCH(c, adj_backward, c, dL_da, dL_dnorm, dL_dEt, dL_dEobs, dL_dsoft);
}

int qnd_set_tuning(qnd_ctx *c, int vpt, int tw_contig, int tw_strided)
{
    if (!c) return QND_EINVAL;
    if (vpt != 0 && vpt != 4 && vpt != 8) return fail(c, QND_EINVAL, "values per thread must be 0, 4 or 8");
    if (tw_contig < 0 || tw_strided < 0 || tw_contig > 256 || tw_strided > 256)
        return fail(c, QND_EINVAL, "lines per workgroup must be in [0, 256]");
    c->vpt_override = vpt;
    c->tw_override[0] = tw_contig;
    c->tw_override[1] = tw_strided;
    return 0;
}

// ---------------------------------------------------------------------
// Field synthesis on the device: Field_vectors (src/parameters.c:899-949) at
// the stage times of Initialize_field_vectors (src/integrators.c:74-129) and
// at t_n, for a batch of instances with their own modes.
// ---------------------------------------------------------------------
struct FieldMode {
    double w, F0, xi, phi, delay, tau, ca, sa;  // omega, amplitude, ellipticity, CEP, delay, duration, rotation
    int env;
};
// stage-time increments of one step: t_0 = t_n + a0, t_{j+1} = t_j + inc[j]
struct StageTimes {
    double a0;
    double inc[20];
};

// Envelope f and its derivative df at ts (time since the mode's delay).
__device__ __forceinline__ void envelope_at(const FieldMode &m, double ts, double &f, double &df)
{
    const double pi = 3.141592653589793;
    if (m.env == QND_ENV_CONSTANT) {
        f = 1.0;
        df = 0.0;
    } else if (m.env == QND_ENV_TRAPZ) {
        // two-cycle sin^4 ramps either side of a flat top
        const double tr = 4.0 * pi / m.w;
        const bool up = ts >= 0.0 && ts <= tr, top = ts > tr && ts < m.tau - tr, dn = ts >= m.tau - tr && ts <= m.tau;
        const double su = sin(0.5 * pi * ts / tr), sd = sin(0.5 * pi * (ts - m.tau) / tr);
        f = (up ? su * su * su * su : 0.0) + (top ? 1.0 : 0.0) + (dn ? sd * sd * sd * sd : 0.0);
        df = pi / tr * ((up ? su * su * sin(pi * ts / tr) : 0.0) + (dn ? sd * sd * sin(pi * (ts - m.tau) / tr) : 0.0));
    } else {
        const bool in = ts >= 0.0 && ts <= m.tau;
        const double sn = sin(pi * ts / m.tau);
        if (m.env == QND_ENV_COS2) {
            f = in ? sn * sn : 0.0;
            df = in ? pi * sin(2.0 * pi * ts / m.tau) / m.tau : 0.0;
        } else {  // QND_ENV_COS4
            f = in ? sn * sn * sn * sn : 0.0;
            df = in ? 4.0 * pi * cos(pi * ts / m.tau) * sn * sn * sn / m.tau : 0.0;
        }
    }
}

// Sum of the modes' A(t), E(t) (x, y; the other components zero).
__device__ __forceinline__ void field_at(const FieldMode *modes, int nmodes, double t, double *A, double *E)
{
    double ax = 0.0, ay = 0.0, ex = 0.0, ey = 0.0;
    for (int k = 0; k < nmodes; ++k) {
        const FieldMode &m = modes[k];
        const double ts = t - m.delay;
        double f, df;
        envelope_at(m, ts, f, df);
        double sa, ca;
        sincos(m.w * ts + m.phi, &sa, &ca);
        // A = F0 f (cos, xi sin)(w t + phi) / w, E = -dA/dt
        const double mex = -m.F0 * (df * ca - f * m.w * sa) / m.w;
        const double mey = -m.F0 * m.xi * (df * sa + f * m.w * ca) / m.w;
        const double max_ = m.F0 * f * ca / m.w, may = m.F0 * m.xi * f * sa / m.w;
        ex += m.ca * mex - m.sa * mey;
        ey += m.sa * mex + m.ca * mey;
        ax += m.ca * max_ - m.sa * may;
        ay += m.sa * max_ + m.ca * may;
    }
    A[0] = ax;
    A[1] = ay;
    E[0] = ex;
    E[1] = ey;
}

// One thread per (instance, step): the s stage rows and the t_n row.
static __global__ void k_field(const FieldMode *modes, int nmodes, int nb, int64_t ntime, int s, int dim, double dt,
                               StageTimes st, double *Et, double *At, double *Eo, double *Ao)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (int64_t)nb * ntime) return;
    const int b = (int)(id / ntime);
    const int64_t i = id - (int64_t)b * ntime;
    const FieldMode *mb = modes + (size_t)b * nmodes;
    const double tn = (double)i * dt;
    double A[2], E[2];
    double t = tn + st.a0;
    for (int j = 0; j < s; ++j) {
        field_at(mb, nmodes, t, A, E);
        const int64_t row = (id * s + j) * dim;
        for (int k = 0; k < dim; ++k) {
            Et[row + k] = k < 2 ? E[k] : 0.0;
            At[row + k] = k < 2 ? A[k] : 0.0;
        }
        if (j < s - 1) t += st.inc[j];
    }
    field_at(mb, nmodes, tn, A, E);
    for (int k = 0; k < dim; ++k) {
        Eo[id * dim + k] = k < 2 ? E[k] : 0.0;
        if (Ao) Ao[id * dim + k] = k < 2 ? A[k] : 0.0;
    }
}

// (Re)allocate the batch tables for nb instances of ntime steps, zeroed.
static int alloc_field_batch(qnd_ctx *c, int nb, int64_t ntime)
{
    const size_t ns = (size_t)nb * ntime * c->s * c->dim, no = (size_t)nb * ntime * c->dim;
    if (c->d_bEt) HIPCK(c, hipFree(c->d_bEt));
    if (c->d_bAt) HIPCK(c, hipFree(c->d_bAt));
    if (c->d_bEo) HIPCK(c, hipFree(c->d_bEo));
    c->d_bEt = c->d_bAt = c->d_bEo = nullptr;
    HIPCK(c, hipMalloc(&c->d_bEt, (ns ? ns : 1) * sizeof(double)));
    HIPCK(c, hipMalloc(&c->d_bAt, (ns ? ns : 1) * sizeof(double)));
    HIPCK(c, hipMalloc(&c->d_bEo, (no ? no : 1) * sizeof(double)));
    HIPCK(c, hipMemset(c->d_bEt, 0, (ns ? ns : 1) * sizeof(double)));
    HIPCK(c, hipMemset(c->d_bAt, 0, (ns ? ns : 1) * sizeof(double)));
    HIPCK(c, hipMemset(c->d_bEo, 0, (no ? no : 1) * sizeof(double)));
    c->bnb = nb;
    c->bntime = ntime;
    c->bfield_s = c->s;
    return 0;
}

int qnd_set_field_modes(qnd_ctx *c, int nb, int64_t ntime, int nmodes, const qnd_mode *modes)
{
    if (!c) return QND_EINVAL;
    if (!c->have_int) return fail(c, QND_ESTATE, "integrator not set (qnd_set_integrator)");
    if (nb < 1 || nb > c->n_slots) return fail(c, QND_EINVAL, "batch of %d outside [1, %d] slots", nb, c->n_slots);
    if (ntime < 0) return fail(c, QND_EINVAL, "ntime < 0");
    if (nmodes < 1 || !modes) return fail(c, QND_EINVAL, "at least one mode per instance");
    std::vector<FieldMode> h((size_t)nb * nmodes);
    for (size_t k = 0; k < h.size(); ++k) {
        const qnd_mode &m = modes[k];
        if (m.envelope < QND_ENV_CONSTANT || m.envelope > QND_ENV_COS4)
            return fail(c, QND_EINVAL, "mode %zu: unknown envelope %d", k, m.envelope);
        if (!(m.wavelength > 0.0)) return fail(c, QND_EINVAL, "mode %zu: wavelength must be > 0", k);
        if (!(m.intensity >= 0.0)) return fail(c, QND_EINVAL, "mode %zu: intensity must be >= 0", k);
        if (m.envelope != QND_ENV_CONSTANT && !(m.duration > 0.0))
            return fail(c, QND_EINVAL, "mode %zu: duration must be > 0", k);
        FieldMode &f = h[k];
        f.w = 45.5637639947958 / m.wavelength;  // src/parameters.c:809
        f.F0 = 5.33803e-9 * sqrt(m.intensity);  // src/parameters.c:808
        f.xi = m.ellipticity;
        f.phi = m.cep;
        f.delay = m.delay;
        f.tau = m.duration;
        f.ca = cos(m.angle);
        f.sa = sin(m.angle);
        if (m.angle == 0.0) {  // exactly the reference's (x, y) components
            f.ca = 1.0;
            f.sa = 0.0;
        }
        f.env = m.envelope;
    }
    StageTimes st{};
    st.a0 = c->as_scaled[0];
    for (int j = 0; j + 1 < c->s; ++j) st.inc[j] = c->as_scaled[2 * j + 1] + c->as_scaled[2 * j + 2];
    int rc;
    if ((rc = alloc_field_batch(c, nb, ntime))) return rc;
    const size_t no = (size_t)nb * ntime * c->dim;
    FieldMode *d_modes = nullptr;
    double *d_Ao = nullptr;
    auto cleanup = [&]() {
        if (d_modes) hipFree(d_modes);
        if (d_Ao) hipFree(d_Ao);
    };
    if (hipMalloc(&d_modes, h.size() * sizeof(FieldMode)) != hipSuccess ||
        (nb == 1 && hipMalloc(&d_Ao, (no ? no : 1) * sizeof(double)) != hipSuccess)) {
        cleanup();
        return fail(c, QND_EHIP, "out of device memory for the field modes");
    }
    if (hipMemcpy(d_modes, h.data(), h.size() * sizeof(FieldMode), hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        return fail(c, QND_EHIP, "copy of the field modes failed");
    }
    const int64_t nthreads = (int64_t)nb * ntime;
    if (nthreads > 0) {
        hipLaunchKernelGGL(k_field, dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0, c->stream, d_modes, nmodes,
                           nb, ntime, c->s, c->dim, c->dt, st, c->d_bEt, c->d_bAt, c->d_bEo, d_Ao);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        cleanup();
        return fail(c, QND_EHIP, "field synthesis kernel: %s", hipGetErrorString(e));
    }
    rc = 0;
    if (nb == 1) {
        // the single-instance tables of qnd_set_field as well
        const size_t ns = (size_t)ntime * c->s * c->dim;
        std::vector<double> Et(ns), At(ns), Eo(no), Ao(no);
        if ((ns && (hipMemcpy(Et.data(), c->d_bEt, ns * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
                    hipMemcpy(At.data(), c->d_bAt, ns * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)) ||
            (no && (hipMemcpy(Eo.data(), c->d_bEo, no * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
                    hipMemcpy(Ao.data(), d_Ao, no * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))) {
            cleanup();
            return fail(c, QND_EHIP, "copy of the field tables failed");
        }
        cleanup();
        return qnd_set_field(c, ntime, Et.data(), At.data(), Eo.data(), Ao.data());
    }
    cleanup();
    return rc;
}

int qnd_get_field_batch(qnd_ctx *c, double *Et, double *At, double *Eobs)
{
    if (!c) return QND_EINVAL;
    if (!c->d_bEt) return fail(c, QND_ESTATE, "no batch field tables (qnd_set_field_batch / qnd_set_field_modes)");
    const size_t ns = (size_t)c->bnb * c->bntime * c->bfield_s * c->dim, no = (size_t)c->bnb * c->bntime * c->dim;
    HIPCK(c, hipStreamSynchronize(c->stream));
    if (Et && ns) HIPCK(c, hipMemcpy(Et, c->d_bEt, ns * sizeof(double), hipMemcpyDeviceToHost));
    if (At && ns) HIPCK(c, hipMemcpy(At, c->d_bAt, ns * sizeof(double), hipMemcpyDeviceToHost));
    if (Eobs && no) HIPCK(c, hipMemcpy(Eobs, c->d_bEo, no * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int qnd_set_field_batch(qnd_ctx *c, int nb, int64_t ntime, const double *Et, const double *At,
                        const double *Eobs)
{
    if (!c || nb < 1 || ntime < 0) return QND_EINVAL;
    if (!c->have_int) return fail(c, QND_ESTATE, "integrator not set (qnd_set_integrator)");
    if (nb > c->n_slots) return fail(c, QND_EINVAL, "batch of %d exceeds the %d slots", nb, c->n_slots);
    const size_t ns = (size_t)nb * ntime * c->s * c->dim, no = (size_t)nb * ntime * c->dim;
    int rc;
    if ((rc = alloc_field_batch(c, nb, ntime))) return rc;
    if (Et && ns) HIPCK(c, hipMemcpy(c->d_bEt, Et, ns * sizeof(double), hipMemcpyHostToDevice));
    if (At && ns) HIPCK(c, hipMemcpy(c->d_bAt, At, ns * sizeof(double), hipMemcpyHostToDevice));
    if (Eobs && no) HIPCK(c, hipMemcpy(c->d_bEo, Eobs, no * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int qnd_real_steps_batch(qnd_ctx *c, int slot0, int nb, int gauge, int64_t step0, int64_t nsteps,
                         const int *codes, int ncodes, double *obs_out)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (nb < 1 || slot0 < 0 || slot0 + nb > c->n_slots)
        return fail(c, QND_EINVAL, "slots [%d, %d) outside the %d slots", slot0, slot0 + nb, c->n_slots);
    if (nb > 1 && nb != c->bnb)
        return fail(c, QND_ESTATE, "batch of %d but the field tables hold %d instances (qnd_set_field_batch)", nb,
                    c->bnb);
    if (gauge != QND_GAUGE_LENGTH && gauge != QND_GAUGE_VELOCITY)
        return fail(c, QND_EINVAL, "unknown gauge %d", gauge);
    if (ncodes > 0 && (!codes || !obs_out)) return fail(c, QND_EINVAL, "observable codes need an output array");
    if (nb == 1) return DISPATCH(c, real_steps, c, slot0, gauge, step0, nsteps, codes, ncodes, obs_out);
    return c->prec == QND_C64
               ? real_steps_impl<float>(c, slot0, nb, gauge, step0, nsteps, codes, ncodes, obs_out)
               : real_steps_impl<double>(c, slot0, nb, gauge, step0, nsteps, codes, ncodes, obs_out);
}

int qnd_group_real_steps(qnd_ctx *const *ctxs, int nctx, int gauge, int64_t step0, int64_t nsteps, const int *codes,
                         int ncodes, double *const *obs_out, int use_graph)
{
    if (!ctxs || nctx < 1 || !ctxs[0]) return fail(nullptr, QND_EINVAL, "empty group");
    qnd_ctx *c0 = ctxs[0];
    if (nctx > QND_GROUP_MAX) return fail(c0, QND_EINVAL, "at most %d contexts in a group", QND_GROUP_MAX);
    if (gauge != QND_GAUGE_LENGTH) return fail(c0, QND_EUNSUPPORTED, "groups run the length gauge");
    if (nsteps < 0) return fail(c0, QND_EINVAL, "nsteps < 0");
    if (ncodes > 8 || (ncodes > 0 && !codes)) return fail(c0, QND_EINVAL, "at most 8 observable codes");
    if (!fused_codes(codes, ncodes)) return fail(c0, QND_EUNSUPPORTED, "groups record dipole_r, dipole_a and norm only");
    for (int k = 0; k < nctx; ++k) {
        qnd_ctx *c = ctxs[k];
        if (!c) return fail(c0, QND_EINVAL, "null context in the group");
        int rc = check_ready(c);
        if (rc) return fail(c0, rc, "member %d: %s", k, c->err.c_str());
        for (int j = 0; j < k; ++j)
            if (ctxs[j] == c) return fail(c0, QND_EINVAL, "context %d appears twice in the group", k);
        if (c->prec != c0->prec || c->dim != 2)
            return fail(c0, QND_EUNSUPPORTED, "group members must be 2-D grids of one precision");
        if (c->x64)
            return fail(c0, QND_EUNSUPPORTED, "group launches have no complex64-storage FP64-arithmetic passes");
        if (c->device != c0->device) return fail(c0, QND_EINVAL, "group members must share one device");
        if (c->s != c0->s || memcmp(c->as_scaled, c0->as_scaled, sizeof c->as_scaled) != 0)
            return fail(c0, QND_EINVAL, "group members must share the integrator and DT");
        if (c->bnb < 1 || c->bfield_s != c->s)
            return fail(c0, QND_ESTATE, "member %d has no batch field tables for this integrator "
                        "(qnd_set_field_batch)", k);
        if (step0 < 0 || step0 + nsteps > c->bntime)
            return fail(c0, QND_EINVAL, "steps [%lld, %lld) outside member %d's field table of %lld steps",
                        (long long)step0, (long long)(step0 + nsteps), k, (long long)c->bntime);
        if (ncodes > 0 && (!obs_out || !obs_out[k])) return fail(c0, QND_EINVAL, "observable output of member %d", k);
    }
    int rc = DISPATCH(c0, group_real_steps, ctxs, nctx, step0, nsteps, codes, ncodes, obs_out, use_graph);
    for (int k = 0; k < nctx; ++k) prof_flush(ctxs[k]);
    return rc;
}

int qnd_final_observables(qnd_ctx *c, int slot, const double *R_mask, const double *pr_max, int Ne, double de,
                          double *ped, double *psi_p)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_slot(c, slot))) return rc;
    if (!R_mask || !pr_max) return fail(c, QND_EINVAL, "R_mask and pr_max are required");
    return DISPATCH(c, final_observables, c, slot, R_mask, pr_max, Ne, de, ped, psi_p);
}

int qnd_event_record(qnd_ctx *c, int idx)
{
    if (!c || idx < 0 || idx >= 16) return QND_EINVAL;
    HIPCK(c, hipEventRecord(c->ev[idx], c->stream));
    return 0;
}

int qnd_event_elapsed(qnd_ctx *c, int i0, int i1, float *ms)
{
    if (!c || i0 < 0 || i0 >= 16 || i1 < 0 || i1 >= 16 || !ms) return QND_EINVAL;
    HIPCK(c, hipEventSynchronize(c->ev[i1]));
    HIPCK(c, hipEventElapsedTime(ms, c->ev[i0], c->ev[i1]));
    return 0;
}

int qnd_synchronize(qnd_ctx *c)
{
    if (!c) return QND_EINVAL;
    HIPCK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int qnd_profile_enable(qnd_ctx *c, int enable)
{
    if (!c) return QND_EINVAL;
    prof_flush(c);
    c->prof = enable != 0;
    for (int k = 0; k < PC_N; ++k) {
        c->prof_count[k] = 0;
        c->prof_ms[k] = 0;
    }
    return 0;
}

int qnd_profile_read(qnd_ctx *c, int64_t *count, double *ms, int nclass)
{
    if (!c) return QND_EINVAL;
    prof_flush(c);
    for (int k = 0; k < nclass && k < PC_N; ++k) {
        count[k] = c->prof_count[k];
        ms[k] = c->prof_ms[k];
    }
    return 0;
}

}  // extern "C"
