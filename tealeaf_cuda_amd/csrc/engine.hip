// engine.hip -- TeaLeaf kernel layer for MI355X (gfx950): the C-ABI of
// include/tea_leaf_engine.h over hand-written HIP kernels.
//
// Replaces init_cuda.cu, tea_leaf_kernel_cuda.cu and the *_kernel_cuda.cu files
// of the CUDA reference.  One chunk, one GPU, one HIP stream.  All arithmetic
// is FP64 in the expression order of the reference kernels (kernel_files/
// tea_leaf_*.cuknl); the build disables fused multiply-add contraction so the
// results equal a plain-C evaluation of the same expressions bit for bit.
//
// Field layout in HBM: every field is one (ny+2h) x pitch array of doubles,
// row-major with x fastest (the reference's THARR2D order).  The interior's
// first column sits on a 128-byte boundary and pitch is a multiple of 16
// doubles, so every wavefront row access is a set of whole cache lines.  All
// fields live in one allocation; kernels index them with one flat index
// i = c + r*pitch on pointers pre-shifted by (lead - h).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/tea_leaf_engine.h"
#include "engine_device.hpp"
#include "engine_solver.hpp"
#include "ws_net.hpp"

using namespace tea;

#define EXPORT extern "C" __attribute__((visibility("default")))

// =============================================================================
// Host state
// =============================================================================
namespace {

enum FieldId {
    F_DENSITY = 1, F_ENERGY0, F_ENERGY1, F_U, F_P, F_SD, F_R,
    X_U0, X_W, X_Z, X_MI, X_KX, X_KY, X_RTEMP, X_UTEMP, X_RSTORE, X_CP, X_BFP, X_VOLUME, X_COUNT,
    X_SD2 = X_COUNT,    // second sd buffer of the fused inner step (internal)
    X_P2,               // second p buffer of the fused outer iteration (internal)
    X_RT2,              // second rtemp buffer of the two-step inner kernel (internal)
    X_U2,               // second u buffer of the fused residual replacement (internal)
    N_SLOTS
};
enum { SOLVER_JACOBI = 0, SOLVER_CG = 1, SOLVER_CHEBY = 2, SOLVER_PPCG = 3 };
// most previous-step differences the warm start combines (tl_warm_start_history)
constexpr int WS_HMAX = 8;
struct WsCoefs {
    double c[WS_HMAX + 1];
};
constexpr int S_CG_FAMILY = (1 << SOLVER_CG) | (1 << SOLVER_CHEBY) | (1 << SOLVER_PPCG);

struct Profiler {
    bool on = false;
    std::string only;   // name prefix of the launches to time (empty: all)
    int every = 1;      // time one in `every` of each name's launches
    struct Pending {
        int name;
        hipEvent_t a, b;
    };
    std::vector<std::string> names;
    std::map<std::string, int> ids;
    std::vector<long> counts;
    std::vector<long> seen;   // launches of each name while wanted
    std::vector<double> ms;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;

    int id(const char *n)
    {
        auto it = ids.find(n);
        if (it != ids.end()) return it->second;
        int k = (int)names.size();
        names.push_back(n);
        ids[n] = k;
        counts.push_back(0);
        seen.push_back(0);
        ms.push_back(0.0);
        return k;
    }
    bool wants(const char *n) const { return on && (only.empty() || std::strncmp(n, only.c_str(), only.size()) == 0); }
    void drain(bool wait);
    hipEvent_t get()
    {
        if (pool.empty() && pending.size() > (1u << 16)) drain(false);   // bound a long profiled run
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
    // Events made ahead, so that a profiled launch costs the host no event
    // creation.
    void reserve(size_t n)
    {
        while (pool.size() < n) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) break;
            pool.push_back(e);
        }
    }
};

// Folds every completed pair into the totals.  Reading events costs the host
// microseconds per pair, so this runs when the profile is read or switched
// off, not at the solver's synchronisation points.
void Profiler::drain(bool wait)
{
    size_t k = 0;
    for (; k < pending.size(); k++) {
        Pending &p = pending[k];
        if (!wait && hipEventQuery(p.b) != hipSuccess) break;
        if (wait) (void)hipEventSynchronize(p.b);
        float t = 0.0f;
        (void)hipEventElapsedTime(&t, p.a, p.b);
        counts[p.name]++;
        ms[p.name] += t;
        pool.push_back(p.a);
        pool.push_back(p.b);
    }
    pending.erase(pending.begin(), pending.begin() + (long)k);
}

struct Chunk {
    bool initialised = false;
    int nx = 0, ny = 0, h = 0;
    int lead = 0;       // physical column of interior column 0
    long pitch = 0;     // row stride in doubles
    long rows = 0;
    int solver = SOLVER_JACOBI;
    int prec = PREC_NONE;
    int device = 0;
    hipStream_t stream = nullptr;

    double *slab = nullptr;         // all 2D fields
    double *f[N_SLOTS] = {nullptr};  // base of each field (physical column 0)
    double *fs[N_SLOTS] = {nullptr}; // shifted: fs[id][c + r*pitch] is cell (c, r)
    double *cellx = nullptr, *celly = nullptr, *vertexx = nullptr, *vertexy = nullptr;
    std::vector<double> h_celldx, h_celldy;

    DD *d_partials = nullptr;        // MAX_RED x max_blocks workgroup partials
    double *h_out = nullptr;         // pinned: the finished values
    double *d_out = nullptr;         // device alias of h_out
    double *d_scal = nullptr;        // finished values kept in device memory for a following launch
    hipEvent_t ev_scal = nullptr;    // marks finished values that later queued launches need not delay
    // rtemp = r and utemp = sd owed by tea_engine_ppcg_begin_ur (0: none, else
    // the field precision in bits): the next streaming inner pass reads r and
    // sd in their place; any other entry point writes them out first.
    int pending_rt_ut = 0;
    // w = A p owed by the fused begin/ur pair (k_ppcg_ur_sd_x2_wp and
    // k_ppcg_ur_sd_wp form w from p instead of reading it), 32 or 64 (the
    // precision of the fields), 0 for none; written out by settle_pending
    int pending_w = 0;
    // z = utemp owed by a separate ppcg_end with a preconditioner (the next
    // fused begin reads utemp in its place and ur rewrites z), 32 or 64, or 0
    int pending_z = 0;
    // z = M^-1 r and sd = z / sd_r_theta (block Jacobi) owed by a fused ur
    // that left them to the first inner pass, which forms sd from r (32 or
    // 64, 0: none); k_minv_r writes them out for any other reader
    int pending_zr = 0, pending_sd_r = 0;
    double sd_r_theta = 0.0;
    // w and r of the u held in the spare u buffer owed by a fused Chebyshev
    // iteration (32 or 64: the precision of the fields, 0: none)
    int pending_cheby = 0;
    // u's halo ring may differ from the spare u buffer's (set by every entry
    // point but the fused Chebyshev iteration, which copies the ring across
    // when it is set)
    bool u_ring_dirty = true;
    // entry points called so far; the end of an outer iteration (update_z,
    // calc_rrn) computed by the last inner pass of a fused call is taken by
    // the next call only when no entry point came in between (end_call: the
    // call that queued it, end_slot: its h_out / d_scal slot, -1: none)
    long calls = 0;
    long end_call = -1;
    int end_slot = -1;
    int max_blocks = 0;              // largest grid of any launch
    int last_grid = 0;               // grid of the latest launch

    double *d_halo_buf = nullptr;    // staging for pack / unpack
    size_t halo_buf_elems = 0;

    std::vector<double> ch_a, ch_b;  // host copies of the uploaded coefficients

    // FP32 shadow fields of the mixed-precision inner solver (allocated on the
    // first tea_engine_set_precision(32)); same cell coordinates, own padding.
    bool fp32 = false;
    float *slab32 = nullptr;
    float *fs32[N_SLOTS] = {nullptr};
    int lead32 = 0;
    long pitch32 = 0;

    // The fused inner step writes the new sd into the spare buffer, interior
    // only; the two buffers' halo rings agree until something writes the sd
    // halo (update_halo, unpack, set_field), after which the ring is copied
    // across before the next fused step.
    bool sd_halo_dirty = false;
    bool sd32_halo_dirty = false;  // the same for the FP32 shadow (set_field32)
    // The same rule for the spare p and rtemp buffers of the fused outer
    // iteration and the two-step inner kernel (index 0: FP64, 1: FP32).
    bool p_halo_dirty[2] = {true, true};
    bool rt_halo_dirty[2] = {true, true};

    // learned warm start: network, inference buffers, u0 of the two previous
    // steps (field layout: the same flat index as the fields)
    bool ws_on = false;
    bool ws_use_net = false;
    ws::Net ws_net;
    ws::Acts ws_acts;
    float *ws_planes = nullptr;   // 3 x ny x nx features
    double *ws_qa = nullptr;      // the network's direction q and A q on the field layout (two fields)
    double *ws_slab = nullptr;
    // previous u0 fields, most recent first (a ring: the oldest buffer takes
    // the next u0), ws_hist of them in use
    double *ws_prev[WS_HMAX + 1] = {nullptr};
    // ws_D[i][j] = v_i . v_j of the history differences v_i = U_i - U_{i+1}
    // (U_0 = u0, U_{i+1} = ws_prev[i]) for i, j < ws_dn: at each step the old
    // v_i become v_{i+1}, so only the row of the new v_0 is reduced
    double ws_D[WS_HMAX + 1][WS_HMAX + 1] = {{0}};
    int ws_dn = 0;
    int ws_hist = 8;              // previous-step differences in the guess (tl_warm_start_history)
    bool ws_minres = true;        // minimise the residual's 2-norm (tl_warm_start_norm residual) or the
                                  // error's energy norm (energy)
    int ws_nprev = 0;             // previous u0 fields held (0..ws_hist + 1)
    double ws_coef[WS_HMAX + 1] = {0};  // coefficients of the last guess (history..., network)

    Profiler prof;
};

Chunk g;
// The deck's profiler_on: per-kernel times are printed when the chunk is
// released or the process ends (the reference prints them in the chunk's
// destructor, cuda_errors.cu:91-106).
bool g_profile_table = false;
std::string g_deck_path = "tea.in";
int g_error_mode = 0;
char g_error[1024] = "";

bool die(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, ap);
    va_end(ap);
    fprintf(stderr, "tea_engine: %s\n", g_error);
    fflush(stderr);
    if (g_error_mode == 0) abort();
    return false;
}

#define HIP_OK(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) { die("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); return; } \
    } while (0)
#define HIP_OK_R(call, ret)                                                                       \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) { die("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); return ret; } \
    } while (0)
void settle_pending();
#define CHECK_INIT_KEEP_PENDING()                                                                 \
    do { if (!g.initialised) { die("%s called before initialise_cuda_", __func__); return; } g.calls++; } while (0)
#define CHECK_INIT()                                                                              \
    do { CHECK_INIT_KEEP_PENDING(); g.u_ring_dirty = true; settle_pending(); } while (0)
// Kernels are registered per solver in the reference (init_cuda.cu:251-282) and
// launching an unregistered one throws there; the same calls are errors here.
#define REQUIRE_SOLVER(mask)                                                                      \
    do { if (!((mask) & (1 << g.solver))) { die("%s: kernel not registered for the solver selected in the deck", __func__); return; } } while (0)

// ---- deck reading: init_cuda.cu:63-163, cuda_strings.cu:7-83 ----------------
std::string match_param(const char *name)
{
    std::ifstream in(g_deck_path);
    std::string line;
    while (std::getline(in, line)) {
        while (!line.empty() && (line.back() == '\r' || line.back() == '\n')) line.pop_back();
        if (line.find('!') != std::string::npos) continue;
        if (line.find(name) != std::string::npos) {
            auto eq = line.find('=');
            return eq != std::string::npos ? line.substr(eq + 1) : std::string(name);
        }
    }
    return std::string();
}
bool param_enabled(const char *name) { return !match_param(name).empty(); }
int read_int(const char *name)
{
    std::string v = match_param(name);
    int x;
    if (v.empty() || sscanf(v.c_str(), " %d", &x) != 1) return -1;
    return x;
}
std::string read_string(const char *name)
{
    std::string v = match_param(name);
    for (auto &ch : v) ch = (char)tolower((unsigned char)ch);
    return v;
}

// ---- launch helpers ------------------------------------------------------------
Ext interior() { return Ext{g.h, g.h + g.nx - 1, g.h, g.h + g.ny - 1}; }

int blocks_for(const Ext &e)
{
    if (e.x1 < e.x0 || e.y1 < e.y0) return 0;
    const int nbx = (e.x1 - e.x0 + TILE_X) / TILE_X;
    const int nby = (e.y1 - e.y0 + TILE_Y) / TILE_Y;
    return nbx * nby;
}

Red red_ctx() { return Red{g.d_partials, g.max_blocks}; }

// The end of a fused FP64 outer iteration (update_z, calc_rrn) rides in its
// last inner pass unless TEA_FUSED_END=0 (FP64 plain PPCG at 4096^2:
// 99.0 -> 97.5 ms per timestep).  Not in FP32: that pass is bound by VALU
// issue, and the compensated FP64 sum it would add cost as much as the
// separate end (45 us) it saves.
template <typename T> bool fused_end()
{
    static bool v = [] {
        const char *s = getenv("TEA_FUSED_END");
        return !(s && s[0] == '0');
    }();
    return v && sizeof(T) == 8;
}

// The PPCG inner step runs as one fused launch unless TEA_FUSED_INNER=0.
bool fused_inner()
{
    static bool v = [] {
        const char *s = getenv("TEA_FUSED_INNER");
        return !(s && s[0] == '0');
    }();
    return v;
}

// One workgroup per tile measured fastest on MI355X for most kernels here (the
// exceptions use launch_capped / launch_x2_capped);
// TEA_GRID_MAX caps the grid (workgroups then loop over several tiles).
int grid_max()
{
    static int v = [] {
        const char *s = getenv("TEA_GRID_MAX");
        int x = s ? atoi(s) : 0;
        return x > 0 ? x : (1 << 30);
    }();
    return v;
}

int grid_for(const Ext &e)
{
    const int nb = blocks_for(e);
    return nb < grid_max() ? nb : grid_max();
}

struct Timed {
    int name = -1;
    hipEvent_t b;
    explicit Timed(const char *n)
    {
        if (!g.prof.wants(n)) return;
        name = g.prof.id(n);
        hipEvent_t a = g.prof.get();
        b = g.prof.get();
        (void)hipEventRecord(a, g.stream);
        g.prof.pending.push_back({name, a, b});
    }
    ~Timed()
    {
        if (name >= 0) (void)hipEventRecord(b, g.stream);
    }
};

// Launches a kernel on the engine's stream.  With the profiler on, the launch
// carries its own start and stop events (hipExtLaunchKernelGGL: the
// timestamps of the dispatch itself), so no marker packets are queued between
// kernels and the profiled run keeps the unprofiled pace.
template <typename K, typename... A>
void klaunch(const char *name, K kernel, dim3 grid, dim3 block, A... args)
{
    if (!g.prof.wants(name)) {
        hipLaunchKernelGGL(kernel, grid, block, 0, g.stream, args...);
        return;
    }
    const int id = g.prof.id(name);
    if (g.prof.seen[id]++ % g.prof.every != 0) {
        hipLaunchKernelGGL(kernel, grid, block, 0, g.stream, args...);
        return;
    }
    hipEvent_t a = g.prof.get(), b = g.prof.get();
    hipExtLaunchKernelGGL(kernel, grid, block, 0, g.stream, a, b, 0u, args...);
    g.prof.pending.push_back({id, a, b});
}

template <typename K, typename... A>
bool launch(const char *name, K kernel, const Ext &e, A... args)
{
    const int nb = grid_for(e);
    g.last_grid = nb;
    if (nb == 0) return true;
    klaunch(name, kernel, dim3(nb), dim3(BLOCK), e, args...);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return die("launch of %s failed: %s", name, hipGetErrorString(err));
    return true;
}

// The same with at most `cap` workgroups, each then looping over several
// tiles (measured faster for a few of the streaming reductions).
template <typename K, typename... A>
bool launch_capped(const char *name, int cap, K kernel, const Ext &e, A... args)
{
    int nb = grid_for(e);
    nb = nb < cap ? nb : cap;
    g.last_grid = nb;
    if (nb == 0) return true;
    klaunch(name, kernel, dim3(nb), dim3(BLOCK), e, args...);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return die("launch of %s failed: %s", name, hipGetErrorString(err));
    return true;
}

// The same for kernels on tiles of 2*TILE_X x TILE_Y cells (FP32 kernels whose
// lanes hold two columns); cap as for launch_capped.
template <typename K, typename... A>
bool launch_x2_capped(const char *name, int cap, K kernel, const Ext &e, A... args)
{
    const int nbx = (e.x1 - e.x0 + 2 * TILE_X) / (2 * TILE_X), nby = (e.y1 - e.y0 + TILE_Y) / TILE_Y;
    int nb = (e.x1 < e.x0 || e.y1 < e.y0) ? 0 : nbx * nby;
    nb = nb < grid_max() ? nb : grid_max();
    nb = nb < cap ? nb : cap;
    g.last_grid = nb;
    if (nb == 0) return true;
    klaunch(name, kernel, dim3(nb), dim3(BLOCK), e, args...);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return die("launch of %s failed: %s", name, hipGetErrorString(err));
    return true;
}

template <typename K, typename... A>
bool launch_x2(const char *name, K kernel, const Ext &e, A... args)
{
    return launch_x2_capped(name, 1 << 30, kernel, e, args...);
}

// Combines the partials of the last reduction launch on the device into
// h_out[slot .. slot+n-1] (and dev_out[0 .. n-1] if given), without waiting.
bool finish(int slot, int n, double *dev_out)
{
    klaunch("finish", k_finish, dim3(1), dim3(FINISH_THREADS), (const DD *)g.d_partials, g.max_blocks, g.last_grid,
            n, g.d_out + slot, dev_out);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return die("launch of k_finish failed: %s", hipGetErrorString(err));
    return true;
}

// Waits for the stream and copies h_out[slot .. slot+n-1] to dst.
bool collect(double *dst, int slot, int n)
{
    hipError_t err = hipStreamSynchronize(g.stream);
    if (err != hipSuccess) return die("hipStreamSynchronize failed: %s", hipGetErrorString(err));
    if (g_profile_table && g.prof.on) g.prof.drain(false);   // so the table at exit needs no HIP call
    for (int j = 0; j < n; j++) dst[j] = g.h_out[slot + j];
    return true;
}

// Combines the partials of the last reduction launch on the device, waits,
// and returns the n finished values.
bool fetch(double *dst, int n) { return finish(0, n, nullptr) && collect(dst, 0, n); }

}  // namespace

// =============================================================================
// Kernels
// =============================================================================
namespace {


// ---- set-up ---------------------------------------------------------------------

// initialise_chunk_kernel.cuknl:54-70 (volume only; areas are not read on this path)
__global__ __launch_bounds__(BLOCK) void k_init_volume(Ext e, long pitch, double *__restrict__ vol, double v)
{
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) vol[i0 + k * pitch] = v;
    STRIP_END
}

// generate_chunk_kernel.cuknl:85-99
__global__ __launch_bounds__(BLOCK) void k_generate_init(Ext e, long pitch, double *__restrict__ dens,
                                                         double *__restrict__ e0, double sd, double se)
{
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        e0[i0 + k * pitch] = se;
        dens[i0 + k * pitch] = sd;
    }
    STRIP_END
}

// generate_chunk_kernel.cuknl:5-67 (one state)
__global__ __launch_bounds__(BLOCK) void k_generate_state(Ext e, long pitch, const double *__restrict__ vertexx,
                                                          const double *__restrict__ vertexy,
                                                          const double *__restrict__ cellx,
                                                          const double *__restrict__ celly,
                                                          double *__restrict__ dens, double *__restrict__ e0,
                                                          double sdens, double senergy, double sxmin, double sxmax,
                                                          double symin, double symax, double sradius, int geom,
                                                          int g_rect, int g_circ, int g_point)
{
    STRIP_BEGIN
    const double x_cent = sxmin, y_cent = symin;
    for (int k = 0; k < s.len; k++) {
        const int c = s.c, r = s.r0 + k;
        bool hit = false;
        if (g_rect == geom) {
            hit = vertexx[1 + c] >= sxmin && vertexx[c] < sxmax && vertexy[1 + r] >= symin && vertexy[r] < symax;
        } else if (geom == g_circ) {
            const double x_pos = cellx[c] - x_cent;
            const double y_pos = celly[r] - y_cent;
            const double radius = sqrt(x_pos * x_pos + y_pos * y_pos);
            hit = radius <= sradius;
        } else if (geom == g_point) {
            hit = vertexx[c] == x_cent && vertexy[r] == y_cent;
        }
        if (hit) {
            e0[i0 + k * pitch] = senergy;
            dens[i0 + k * pitch] = sdens;
        }
    }
    STRIP_END
}

// generate_chunk_kernel.cuknl:69-83: u = u0 = energy*density
__global__ __launch_bounds__(BLOCK) void k_init_u(Ext e, long pitch, const double *__restrict__ dens,
                                                  const double *__restrict__ energy, double *__restrict__ u,
                                                  double *__restrict__ u0)
{
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        u[i] = energy[i] * dens[i];
        u0[i] = energy[i] * dens[i];
    }
    STRIP_END
}

// field_summary_kernel.cuknl:5-58
__global__ __launch_bounds__(BLOCK) void k_field_summary(Ext e, long pitch, const double *__restrict__ vol,
                                                         const double *__restrict__ dens,
                                                         const double *__restrict__ energy,
                                                         const double *__restrict__ u, Red red)
{
    DD acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        const double cell_vol = vol[i];
        const double cell_mass = cell_vol * dens[i];
        dd_add_prod(acc[0], cell_vol, 1.0);
        dd_add_prod(acc[1], cell_vol, dens[i]);
        dd_add_prod(acc[2], cell_mass, energy[i]);
        dd_add_prod(acc[3], cell_mass, u[i]);
    }
    STRIP_END
    block_reduce<4>(acc, red);
}

// ---- halos: reflective update and pack / unpack ------------------------------------
// One thread per (position along the face, depth level).  Reflection: the halo
// cell at distance k from the face takes the interior cell at distance k-1.

__global__ void k_reflect_bt(double *__restrict__ a, long pitch, int h, int nx, int ny, int d, int top)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;  // column offset in [0, nx+2d)
    const int k = blockIdx.y + 1;                         // distance 1..d
    if (j >= nx + 2 * d) return;
    const int c = h - d + j;
    if (!top) a[c + (long)(h - k) * pitch] = a[c + (long)(h + k - 1) * pitch];
    else a[c + (long)(h + ny - 1 + k) * pitch] = a[c + (long)(h + ny - k) * pitch];
}

__global__ void k_reflect_lr(double *__restrict__ a, long pitch, int h, int nx, int ny, int d, int right)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;  // row offset in [0, ny+2d)
    const int k = blockIdx.y + 1;
    if (j >= ny + 2 * d) return;
    const long r = h - d + j;
    if (!right) a[(h - k) + r * pitch] = a[(h + k - 1) + r * pitch];
    else a[(h + nx - 1 + k) + r * pitch] = a[(h + nx - k) + r * pitch];
}

// Copies the halo ring (every cell outside the interior) of src into dst:
// bottom rows, top rows, then the left and right columns of the interior rows.
template <typename T>
__global__ void k_copy_ring(T *__restrict__ dst, const T *__restrict__ src, long pitch, int h, int nx, int ny)
{
    const long w = nx + 2L * h;
    const long n_bt = w * h;           // cells in the bottom (or top) band
    const long n_lr = (long)ny * h;    // cells in the left (or right) band
    long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long c, r;
    if (j < n_bt) { c = j % w; r = j / w; }
    else if ((j -= n_bt) < n_bt) { c = j % w; r = h + ny + j / w; }
    else if ((j -= n_bt) < n_lr) { c = j % h; r = h + j / h; }
    else if ((j -= n_lr) < n_lr) { c = h + nx + j % h; r = h + j / h; }
    else return;
    dst[c + r * pitch] = src[c + r * pitch];
}

// pack_kernel.cuknl:5-147 buffer layout: left/right strips are (ny+2d) x d with
// element (j,k) at j + k*d; bottom/top strips are d x (nx+2d) with (j,k) at
// j + k*(nx+2d).  Left and bottom strips hold the face cells mirrored.
__global__ void k_pack(double *__restrict__ a, double *__restrict__ buf, long pitch, int h, int nx, int ny, int d,
                       int face, int pack)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    long ai, bi;
    if (face == TEA_CHUNK_LEFT || face == TEA_CHUNK_RIGHT) {
        // j runs over rows (ny+2d), k over the d columns
        if (j >= ny + 2 * d) return;
        const long r = h - d + j;
        int c;
        if (face == TEA_CHUNK_LEFT) c = pack ? (h + d - 1 - k) : (h - d + k);
        else c = pack ? (h + nx - d + k) : (h + nx + d - 1 - k);
        ai = c + r * pitch;
        bi = k + (long)j * d;
    } else {
        if (j >= nx + 2 * d) return;
        const int c = h - d + j;
        long r;
        if (face == TEA_CHUNK_BOTTOM) r = pack ? (h + d - 1 - k) : (h - d + k);
        else r = pack ? (h + ny - d + k) : (h + ny + d - 1 - k);
        ai = c + r * pitch;
        bi = j + (long)k * (nx + 2 * d);
    }
    if (pack) buf[bi] = a[ai];
    else a[ai] = buf[bi];
}

// ---- solver: common --------------------------------------------------------------

// tea_leaf_common.cuknl:60-97 over x,y in [0,+h]
__global__ __launch_bounds__(BLOCK) void k_init_common(Ext e, long pitch, const double *__restrict__ dens,
                                                       double *__restrict__ Kx, double *__restrict__ Ky,
                                                       double rx, double ry, int coef, long i_origin)
{
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        double dc = 0, dl = 0, dd = 0;
        if (coef == TEA_CONDUCTIVITY) {
            dc = dens[i];
            if (i - 1 - i_origin > 0) dl = dens[i - 1];
            if (i - pitch - i_origin > 0) dd = dens[i - pitch];
        } else {
            dc = 1.0 / dens[i];
            dl = 1.0 / dens[i - 1];
            dd = 1.0 / dens[i - pitch];
        }
        double kx = (dl + dc) / (2.0 * dl * dc);
        kx *= rx;
        double ky = (dd + dc) / (2.0 * dd * dc);
        ky *= ry;
        Kx[i] = kx;
        Ky[i] = ky;
    }
    STRIP_END
}

// tea_leaf_common.cuknl:99-129 over +-h
__global__ __launch_bounds__(BLOCK) void k_zero_boundaries(Ext e, long pitch, double *__restrict__ Kx,
                                                           double *__restrict__ Ky, int h, int nx, int ny, int zl,
                                                           int zr, int zb, int zt)
{
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        const int c = s.c, r = s.r0 + k;
        if (zl && c <= h) Kx[i] = 0;
        if (zr && c > (nx - 1) + h) Kx[i] = 0;
        if (zb && r <= h) Ky[i] = 0;
        if (zt && r > (ny - 1) + h) Ky[i] = 0;
    }
    STRIP_END
}

// tea_leaf_common.cuknl:19-39: r = u0 - A u
__global__ __launch_bounds__(BLOCK) void k_calc_residual(Ext e, long pitch, const double *__restrict__ u,
                                                         const double *__restrict__ u0, double *__restrict__ r,
                                                         const double *__restrict__ Kx,
                                                         const double *__restrict__ Ky)
{
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const double smvp = apply_A(u, Kx, Ky, i, pitch);
            r[i] = u0[i] - smvp;
        }
    }
    STRIP_END
}

// tea_leaf_common.cuknl:5-17: energy1 = u/density
__global__ __launch_bounds__(BLOCK) void k_finalise(Ext e, long pitch, const double *__restrict__ dens,
                                                    const double *__restrict__ u, double *__restrict__ e1)
{
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) e1[i0 + k * pitch] = u[i0 + k * pitch] / dens[i0 + k * pitch];
    STRIP_END
}

// ---- learned warm start (no reference counterpart) ---------------------------------

// sums of (u0 - u_prev)^2 and u0^2 over the interior
__global__ __launch_bounds__(BLOCK) void k_ws_stats(Ext e, long pitch, const double *__restrict__ u0,
                                                    const double *__restrict__ up, Red red)
{
    DD acc[2] = {{0, 0}, {0, 0}};
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        const double d = u0[i] - up[i];
        dd_add_prod(acc[0], d, d);
        dd_add_prod(acc[1], u0[i], u0[i]);
    }
    STRIP_END
    block_reduce<2>(acc, red);
}

// network features on dense planes [3][ny][nx]
__global__ __launch_bounds__(BLOCK) void k_ws_features(Ext e, long pitch, int h, int nx, long n,
                                                       const double *__restrict__ dens, const double *__restrict__ u0,
                                                       const double *__restrict__ up, double inv_s, double inv_su,
                                                       float *__restrict__ planes)
{
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        const long o = (long)(s.r0 + k - h) * nx + (s.c - h);
        planes[o] = (float)((log10(dens[i]) - 0.5) / 1.5);
        planes[n + o] = (float)((u0[i] - up[i]) * inv_s);
        planes[2 * n + o] = (float)(u0[i] * inv_su);
    }
    STRIP_END
}

// Search directions of the warm start at interior cell i: the differences of
// the previous steps v_j = up_j - up_{j+1} (up_0 = u0, up_1 = the previous
// step's u0, ...) for j < nh, and v_HN = scale * net (when net != nullptr);
// the network plane is dense, its cells outside the interior count as 0
// (their faces carry K = 0 under the zero-flux boundary).
struct WsDirs {
    const double *u0;
    const double *up[WS_HMAX + 1];
    const double *la, *lb;   // has_last: the two fields of the last direction (U_{nd-1}, U_nd)
    const float *net;
    double scale;
    const double *q, *aq;   // the network's direction (scale * net) and A q, from k_ws_netdir
    int h, nx, ny;
    int nh;   // history directions in the guess (0..HN)
    int nv;   // differences v_0..v_{nv-1} the Gram pass reduces against v_0 (nv <= HN + 1)
    int has_last;   // A v_{nh-1} by the stencil (v_nh not held yet)
};

__device__ __forceinline__ double ws_net_at(const WsDirs &d, int c, int r)
{
    if (c < d.h || c >= d.h + d.nx || r < d.h || r >= d.h + d.ny) return 0.0;
    return d.scale * (double)d.net[(long)(r - d.h) * d.nx + (c - d.h)];
}

// Combines per-lane sums over the workgroup in compensated arithmetic (each
// lane's sum covers the few cells of its strips) and publishes the
// workgroup's partials; one value's wave reduction at a time, so the lanes
// keep only their plain sums.
template <int NV>
__device__ __forceinline__ void block_reduce_lanes(const double (&v)[NV], const Red red)
{
    __shared__ DD lds[NV][WAVES];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
#pragma unroll
    for (int j = 0; j < NV; j++) {
        const DD t = wave_reduce(DD{v[j], 0.0});
        if (lane == 0) lds[j][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 0; j < NV; j++) {
            DD t = lds[j][0];
            for (int w = 1; w < WAVES; w++) dd_merge(t, lds[j][w]);
            red.partials[j * red.max_blocks + blockIdx.x] = t;
        }
    }
}

// The new dot products of the warm start's Gram system at this step (the host
// keeps the rest from the steps before).  Each earlier step's solve left
// A U_a = U_{a+1} to its tolerance (U_a was that step's solution and U_{a+1}
// its right-hand side), so for the differences v_a = U_a - U_{a+1},
// A v_a = v_{a+1}, and the residual of the u = u0 guess is r0 = v_0: every
// entry of the system is a dot product v_i . v_j, and only v_0 is new.  Out
// (in this order): v_0 . v_j for j < nv; with has_last (A v_{nh-1} by the
// stencil on la - lb, while v_nh is not held yet) last . last and last . v_j
// (j < nv); with the network's direction q (A q by the stencil) A q . A q,
// q . A q, q . v_0, A q . v_j (j < nv) and, with has_last, A q . last.
// q = scale * net on the field layout and A q (the stencil over q, zero
// outside the interior), for the Gram pass and the guess.
__global__ __launch_bounds__(BLOCK) void k_ws_netdir(Ext e, long pitch, WsDirs d, const double *__restrict__ Kx,
                                                     const double *__restrict__ Ky, double *__restrict__ q,
                                                     double *__restrict__ aq)
{
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        const int c = s.c, r = s.r0 + k;
        const double qv = ws_net_at(d, c, r);
        const double ky1 = Ky[i + pitch], ky0 = Ky[i], kx1 = Kx[i + 1], kx0 = Kx[i];
        q[i] = qv;
        aq[i] = (1.0 + (ky1 + ky0) + (kx1 + kx0)) * qv -
                (ky1 * ws_net_at(d, c, r + 1) + ky0 * ws_net_at(d, c, r - 1)) -
                (kx1 * ws_net_at(d, c + 1, r) + kx0 * ws_net_at(d, c - 1, r));
    }
    STRIP_END
}

// LAST and NET (has_last and the network) are template switches: the
// steady-state pass without the network keeps NV sums only.
template <int HN, bool LAST, bool NET>
__global__ __launch_bounds__(BLOCK) void k_ws_gram(Ext e, long pitch, WsDirs d, const double *__restrict__ Kx,
                                                   const double *__restrict__ Ky, Red red)
{
    constexpr int NV = HN + 1, NACC = NET ? 3 * NV + 5 : (LAST ? 2 * NV + 1 : NV);
    const int nv = d.nv;
    double acc[NACC];
#pragma unroll
    for (int j = 0; j < NACC; j++) acc[j] = 0.0;
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        const int c = s.c, r = s.r0 + k;
        double v[NV];
        double prev = d.u0[i];
#pragma unroll
        for (int j = 0; j < NV; j++) {
            v[j] = 0.0;
            if (j < nv) {
                const double uj = d.up[j][i];
                v[j] = prev - uj;
                prev = uj;
            }
        }
#pragma unroll
        for (int j = 0; j < NV; j++) acc[j] += v[0] * v[j];
        constexpr int q0 = NV, q = 2 * NV + 1;
        double last = 0.0;
        if constexpr (LAST || NET) {
            if constexpr (LAST) {
                const double ky1 = Ky[i + pitch], ky0 = Ky[i], kx1 = Kx[i + 1], kx0 = Kx[i];
                last = (1.0 + (ky1 + ky0) + (kx1 + kx0)) * (d.la[i] - d.lb[i]) -
                       (ky1 * (d.la[i + pitch] - d.lb[i + pitch]) + ky0 * (d.la[i - pitch] - d.lb[i - pitch])) -
                       (kx1 * (d.la[i + 1] - d.lb[i + 1]) + kx0 * (d.la[i - 1] - d.lb[i - 1]));
                acc[q0] += last * last;
#pragma unroll
                for (int j = 0; j < NV; j++) acc[q0 + 1 + j] += last * v[j];
            }
            if constexpr (NET) {
                const double qv = d.q[i], aq = d.aq[i];
                acc[q] += aq * aq;
                acc[q + 1] += qv * aq;
                acc[q + 2] += qv * v[0];
                acc[q + 3] += aq * last;
#pragma unroll
                for (int j = 0; j < NV; j++) acc[q + 4 + j] += aq * v[j];
            }
        }
        (void)c;
        (void)r;
    }
    STRIP_END
    block_reduce_lanes<NACC>(acc, red);
}

// u = u0 + sum_j c[j] v_j + c[HN] v_HN (apply == 0: no guess); and u0 goes to
// `keep`, the buffer of the oldest previous field read here (the host then
// rotates the ring: keep becomes up_1).
template <int HN>
__global__ __launch_bounds__(BLOCK) void k_ws_apply(Ext e, long pitch, double *__restrict__ u, WsDirs d,
                                                    double *__restrict__ keep, WsCoefs cf, int apply)
{
    STRIP_BEGIN
    for (int k = 0; k < s.len; k++) {
        const long i = i0 + k * pitch;
        const double u0 = d.u0[i];
        if (apply) {
            double v = u0 + cf.c[0] * (u0 - d.up[0][i]);
            double up_prev = d.up[0][i];
#pragma unroll
            for (int j = 1; j < HN; j++) {
                if (j < d.nh && cf.c[j] != 0.0) v += cf.c[j] * (up_prev - d.up[j][i]);
                if (j < d.nh) up_prev = d.up[j][i];
            }
            if (cf.c[HN] != 0.0) v += cf.c[HN] * d.q[i];   // q = scale * net (k_ws_netdir)
            u[i] = v;
        }
        keep[i] = u0;
    }
    STRIP_END
}

// ---- solver: CG -------------------------------------------------------------------

// ---- solver: Chebyshev --------------------------------------------------------------

// tea_leaf_cheby.cuknl:5-68 (first = 1, p = M^-1 r / theta) and :84-152
// (p = alpha p + beta M^-1 r); both start with w = A u, r = u0 - w.  T =
// float is the inner solver of the mixed-precision refinement (scalars
// rounded to T once).
template <typename T, int PREC, int FIRST>
__global__ __launch_bounds__(BLOCK) void k_cheby_p(Ext e, long pitch, const T *__restrict__ u,
                                                   const T *__restrict__ u0, T *__restrict__ p, T *__restrict__ r,
                                                   T *__restrict__ w, const T *__restrict__ Mi,
                                                   const T *__restrict__ Kx, const T *__restrict__ Ky, double theta_in,
                                                   double alpha_in, double beta_in)
{
    const T theta = (T)theta_in, alpha = (T)alpha_in, beta = (T)beta_in;
    STRIP_BEGIN
    T rv[STRIP] = {0, 0, 0, 0}, zv[STRIP] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const T wv = apply_A(u, Kx, Ky, i, pitch);
            w[i] = wv;
            rv[k] = u0[i] - wv;
            r[i] = rv[k];
        }
    }
    if (PREC == PREC_BLOCK) precond_strip<T, PREC_BLOCK>(rv, zv, s.len, i0, pitch, Mi, Kx, Ky);
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            T pv;
            if (FIRST) {
                if (PREC == PREC_BLOCK) pv = zv[k] / theta;
                else if (PREC == PREC_DIAG) pv = (Mi[i] * rv[k]) / theta;
                else pv = rv[k] / theta;
            } else {
                if (PREC == PREC_BLOCK) pv = alpha * p[i] + beta * zv[k];
                else if (PREC == PREC_DIAG) pv = alpha * p[i] + beta * Mi[i] * rv[k];
                else pv = alpha * p[i] + beta * rv[k];
            }
            p[i] = pv;
        }
    }
    STRIP_END
}

// tea_leaf_cheby.cuknl:71-82: u += p
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_add(Ext e, long pitch, T *__restrict__ u, const T *__restrict__ p)
{
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++)
        if (k < s.len) u[i0 + k * pitch] += p[i0 + k * pitch];
    STRIP_END
}

// ---- solver: Jacobi -------------------------------------------------------------------

// tea_leaf_jacobi.cuknl:18-46 (un is the copy of u held in the Mi buffer)
__global__ __launch_bounds__(BLOCK) void k_jacobi(Ext e, long pitch, const double *__restrict__ Kx,
                                                  const double *__restrict__ Ky, const double *__restrict__ u0,
                                                  double *__restrict__ u, const double *__restrict__ un, Red red)
{
    DD acc[1] = {{0, 0}};
    STRIP_BEGIN
#pragma unroll
    for (int k = 0; k < STRIP; k++) {
        if (k < s.len) {
            const long i = i0 + k * pitch;
            const double uv = (u0[i] + Kx[i + 1] * un[i + 1] + Kx[i] * un[i - 1] + Ky[i + pitch] * un[i + pitch]
                               + Ky[i] * un[i - pitch])
                              / (1.0 + (Kx[i] + Kx[i + 1]) + (Ky[i] + Ky[i + pitch]));
            u[i] = uv;
            dd_add_prod(acc[0], fabs(uv - un[i]), 1.0);
        }
    }
    STRIP_END
    block_reduce<1>(acc, red);
}

// ---- solver: PPCG ---------------------------------------------------------------------

}  // namespace

// =============================================================================
// Entry points
// =============================================================================
namespace {

double *F(int id) { return g.fs[id]; }

// The profiler_on table: total milliseconds per kernel, by name, in the
// reference's format (cuda_errors.cu:95-105).  Folds in the completed launches
// first when the stream still exists.
void print_profile_table(bool hip_alive)
{
    if (!g_profile_table) return;
    g_profile_table = false;
    if (hip_alive && g.stream) {
        (void)hipStreamSynchronize(g.stream);
        g.prof.drain(false);
    }
    std::map<std::string, double> byname;
    for (size_t k = 0; k < g.prof.names.size(); k++)
        if (g.prof.counts[k] > 0) byname[g.prof.names[k]] += g.prof.ms[k];
    fprintf(stdout, "@@@@@ PROFILING @@@@@\n");
    for (const auto &kv : byname) fprintf(stdout, "%35s : %.3f\n", kv.first.c_str(), kv.second);
    fflush(stdout);
}

// At process exit: the table of a chunk that was never released (no HIP calls
// then; the launches were folded in at the last synchronisation).
struct ProfileAtExit {
    ~ProfileAtExit() { print_profile_table(false); }
} g_profile_at_exit;

void release_all()
{
    print_profile_table(true);
    if (g.stream) (void)hipStreamSynchronize(g.stream);
    for (auto &p : g.prof.pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    g.prof.pending.clear();
    for (auto e : g.prof.pool) (void)hipEventDestroy(e);
    g.prof.pool.clear();
    if (g.slab) (void)hipFree(g.slab);
    if (g.cellx) (void)hipFree(g.cellx);
    if (g.d_partials) (void)hipFree(g.d_partials);
    if (g.h_out) (void)hipHostFree(g.h_out);
    if (g.d_scal) (void)hipFree(g.d_scal);
    if (g.ev_scal) (void)hipEventDestroy(g.ev_scal);
    g.ev_scal = nullptr;
    if (g.d_halo_buf) (void)hipFree(g.d_halo_buf);
    if (g.slab32) (void)hipFree(g.slab32);
    ws::net_free(g.ws_net);
    ws::acts_free(g.ws_acts);
    if (g.ws_planes) (void)hipFree(g.ws_planes);
    if (g.ws_qa) (void)hipFree(g.ws_qa);
    if (g.ws_slab) (void)hipFree(g.ws_slab);
    if (g.stream) (void)hipStreamDestroy(g.stream);
    Profiler keep = g.prof;
    g = Chunk();
    g.prof.on = keep.on;
}


// Dispatch a preconditioner-templated kernel.
#define PREC_SWITCH(PREC_VALUE, KERNEL, ...)                                    \
    ((PREC_VALUE) == PREC_BLOCK ? launch(#KERNEL, KERNEL<PREC_BLOCK>, __VA_ARGS__) \
     : (PREC_VALUE) == PREC_DIAG ? launch(#KERNEL, KERNEL<PREC_DIAG>, __VA_ARGS__) \
                                 : launch(#KERNEL, KERNEL<PREC_NONE>, __VA_ARGS__))

}  // namespace

EXPORT void initialise_cuda_(int *x_min, int *x_max, int *y_min, int *y_max)
{
    release_all();
    {
        std::ifstream probe(g_deck_path);
        if (!probe) { die("Input file not found: %s", g_deck_path.c_str()); return; }
    }
    int dev = read_int("hip_device");
    if (dev < 0) dev = read_int("cuda_device");
    if (dev < 0) dev = 0;
    int ndev = 0;
    HIP_OK(hipGetDeviceCount(&ndev));
    if (dev >= ndev) { die("Device id %d specified in the deck, but only %d devices available", dev, ndev); return; }
    HIP_OK(hipSetDevice(dev));
    g.device = dev;

    g.h = read_int("halo_depth");
    if (g.h < 1) { die("Halo exchange depth unspecified or was too small"); return; }
    if (param_enabled("tl_use_ppcg")) g.solver = SOLVER_PPCG;
    else if (param_enabled("tl_use_chebyshev")) g.solver = SOLVER_CHEBY;
    else if (param_enabled("tl_use_cg")) g.solver = SOLVER_CG;
    else g.solver = SOLVER_JACOBI;
    const std::string prec = read_string("tl_preconditioner_type");
    if (prec.find("jac_diag") != std::string::npos) g.prec = PREC_DIAG;
    else if (prec.find("jac_block") != std::string::npos) g.prec = PREC_BLOCK;
    else g.prec = PREC_NONE;
    if (param_enabled("profiler_on")) {
        g.prof.on = true;
        g_profile_table = true;
    }

    g.nx = *x_max - *x_min + 1;
    g.ny = *y_max - *y_min + 1;
    if (g.nx < 1 || g.ny < 1) { die("invalid chunk size %d x %d", g.nx, g.ny); return; }
    g.lead = (g.h + 15) / 16 * 16;
    g.pitch = (long)((g.lead + g.nx + g.h + 15) / 16 * 16);
    g.rows = (long)g.ny + 2L * g.h;

    HIP_OK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    const size_t field_elems = (size_t)g.pitch * (size_t)g.rows + 16;  // + slack, keeps 128-byte alignment
    const size_t field_stride = (field_elems + 31) / 32 * 32;
    HIP_OK(hipMalloc(&g.slab, field_stride * N_SLOTS * sizeof(double)));
    HIP_OK(hipMemsetAsync(g.slab, 0, field_stride * N_SLOTS * sizeof(double), g.stream));
    for (int k = 1; k < N_SLOTS; k++) {
        g.f[k] = g.slab + field_stride * k;
        g.fs[k] = g.f[k] + (g.lead - g.h);
    }
    // 1D coordinate arrays (cell and vertex x/y), indices 0 .. n+2h
    const size_t nxv = (size_t)g.nx + 2 * g.h + 2, nyv = (size_t)g.ny + 2 * g.h + 2;
    HIP_OK(hipMalloc(&g.cellx, (2 * nxv + 2 * nyv) * sizeof(double)));
    g.vertexx = g.cellx + nxv;
    g.celly = g.vertexx + nxv;
    g.vertexy = g.celly + nyv;

    {
        const int all = blocks_for(Ext{0, g.nx + 2 * g.h - 1, 0, g.ny + 2 * g.h - 1});
        g.max_blocks = (all < grid_max() ? all : grid_max()) + 1;
    }
    HIP_OK(hipMalloc(&g.d_partials, (size_t)MAX_RED * g.max_blocks * sizeof(DD)));
    HIP_OK(hipHostMalloc(&g.h_out, MAX_RED * sizeof(double), hipHostMallocDefault));
    HIP_OK(hipHostGetDevicePointer((void **)&g.d_out, g.h_out, 0));
    HIP_OK(hipMalloc(&g.d_scal, MAX_RED * sizeof(double)));
    if (!g.ev_scal) HIP_OK(hipEventCreateWithFlags(&g.ev_scal, hipEventDisableTiming));

    g.halo_buf_elems = (size_t)TEA_NUM_FIELDS * (size_t)g.h * (size_t)((g.nx > g.ny ? g.nx : g.ny) + 2 * g.h);
    HIP_OK(hipMalloc(&g.d_halo_buf, g.halo_buf_elems * sizeof(double)));
    HIP_OK(hipStreamSynchronize(g.stream));
    g.initialised = true;
}

EXPORT void initialise_chunk_kernel_cuda_(double *xmin, double *ymin, double *dx, double *dy)
{
    CHECK_INIT();
    const double d_xmin = *xmin, d_ymin = *ymin, d_dx = *dx, d_dy = *dy;
    const int h = g.h;
    const size_t nxv = (size_t)g.nx + 2 * h + 2, nyv = (size_t)g.ny + 2 * h + 2;
    std::vector<double> host(2 * nxv + 2 * nyv, 0.0);
    double *cx = host.data(), *vx = cx + nxv, *cy = vx + nxv, *vy = cy + nyv;
    g.h_celldx.assign(nxv, 0.0);
    g.h_celldy.assign(nyv, 0.0);
    // initialise_chunk_kernel.cuknl:3-51
    for (int c = 0; c <= g.nx + 2 * h; c++) {
        vx[c] = d_xmin + d_dx * (double)((c - 1) - h + 1);
        const double vplus = d_xmin + d_dx * (double)(c - h + 1);
        cx[c] = 0.5 * (vx[c] + vplus);
        g.h_celldx[c] = d_dx;
    }
    for (int r = 0; r <= g.ny + 2 * h; r++) {
        vy[r] = d_ymin + d_dy * (double)((r - 1) - h + 1);
        const double vplus = d_ymin + d_dy * (double)(r - h + 1);
        cy[r] = 0.5 * (vy[r] + vplus);
        g.h_celldy[r] = d_dy;
    }
    HIP_OK(hipMemcpyAsync(g.cellx, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice, g.stream));
    const Ext all{0, g.nx + 2 * h - 1, 0, g.ny + 2 * h - 1};
    launch("initialise_chunk", k_init_volume, all, g.pitch, F(X_VOLUME), d_dx * d_dy);
    HIP_OK(hipStreamSynchronize(g.stream));
}

EXPORT void generate_chunk_kernel_cuda_(const int *number_of_states, const double *state_density,
                                        const double *state_energy, const double *state_xmin,
                                        const double *state_xmax, const double *state_ymin,
                                        const double *state_ymax, const double *state_radius,
                                        const int *state_geometry, const int *g_rect, const int *g_circ,
                                        const int *g_point)
{
    CHECK_INIT();
    const int ns = *number_of_states;
    if (ns < 1) { die("generate_chunk: no states"); return; }
    const Ext in = interior();
    launch("generate_chunk_init", k_generate_init, in, g.pitch, F(F_DENSITY), F(F_ENERGY0), state_density[0],
           state_energy[0]);
    for (int s = 1; s < ns; s++) {
        launch("generate_chunk_kernel", k_generate_state, in, g.pitch, (const double *)g.vertexx,
               (const double *)g.vertexy, (const double *)g.cellx, (const double *)g.celly, F(F_DENSITY),
               F(F_ENERGY0), state_density[s], state_energy[s], state_xmin[s], state_xmax[s], state_ymin[s],
               state_ymax[s], state_radius[s], state_geometry[s], *g_rect, *g_circ, *g_point);
    }
    launch("generate_chunk_init_u", k_init_u, in, g.pitch, (const double *)F(F_DENSITY),
           (const double *)F(F_ENERGY0), F(F_U), F(X_U0));
    HIP_OK(hipStreamSynchronize(g.stream));
}

EXPORT void set_field_kernel_cuda_(void)
{
    CHECK_INIT();
    launch("set_field", k_copy<double>, interior(), g.pitch, F(F_ENERGY1), (const double *)F(F_ENERGY0));
}

EXPORT void field_summary_kernel_cuda_(double *vol, double *mass, double *ie, double *temp)
{
    CHECK_INIT();
    if (!launch("field_summary", k_field_summary, interior(), g.pitch, (const double *)F(X_VOLUME),
                (const double *)F(F_DENSITY), (const double *)F(F_ENERGY1), (const double *)F(F_U), red_ctx()))
        return;
    double out[4];
    if (!fetch(out, 4)) return;
    *vol = out[0];
    *mass = out[1];
    *ie = out[2];
    *temp = out[3];
}

EXPORT void update_halo_kernel_cuda_(const int *chunk_neighbours, const int *fields, const int *depth)
{
    CHECK_INIT();
    const int d = *depth;
    if (d < 1 || d > g.h) { die("update_halo: invalid depth %d (halo depth %d)", d, g.h); return; }
    if (fields[TEA_FIELD_SD - 1] == 1) g.sd_halo_dirty = true;
    if (fields[TEA_FIELD_P - 1] == 1) g.p_halo_dirty[0] = true;
    for (int f = 1; f <= TEA_NUM_FIELDS; f++) {
        if (fields[f - 1] != 1) continue;
        double *a = F(f);
        Timed t("update_halo");
        if (chunk_neighbours[TEA_CHUNK_BOTTOM - 1] == TEA_EXTERNAL_FACE)
            hipLaunchKernelGGL(k_reflect_bt, dim3((g.nx + 2 * d + 255) / 256, d), dim3(256), 0, g.stream, a, g.pitch,
                               g.h, g.nx, g.ny, d, 0);
        if (chunk_neighbours[TEA_CHUNK_TOP - 1] == TEA_EXTERNAL_FACE)
            hipLaunchKernelGGL(k_reflect_bt, dim3((g.nx + 2 * d + 255) / 256, d), dim3(256), 0, g.stream, a, g.pitch,
                               g.h, g.nx, g.ny, d, 1);
        if (chunk_neighbours[TEA_CHUNK_LEFT - 1] == TEA_EXTERNAL_FACE)
            hipLaunchKernelGGL(k_reflect_lr, dim3((g.ny + 2 * d + 255) / 256, d), dim3(256), 0, g.stream, a, g.pitch,
                               g.h, g.nx, g.ny, d, 0);
        if (chunk_neighbours[TEA_CHUNK_RIGHT - 1] == TEA_EXTERNAL_FACE)
            hipLaunchKernelGGL(k_reflect_lr, dim3((g.ny + 2 * d + 255) / 256, d), dim3(256), 0, g.stream, a, g.pitch,
                               g.h, g.nx, g.ny, d, 1);
    }
    HIP_OK(hipGetLastError());
}

namespace {
void pack_unpack(int *fields, int *offsets, int d, int face, int pack, double *host_buffer)
{
    CHECK_INIT();
    if (d < 1 || d > g.h) { die("pack: invalid depth %d (halo depth %d)", d, g.h); return; }
    if (face < TEA_CHUNK_LEFT || face > TEA_CHUNK_TOP) {
        die("Invalid face identifier %d passed to mpi buffer packing", face);
        return;
    }
    const bool lr = face <= TEA_CHUNK_RIGHT;
    const long side = lr ? (long)d * (g.ny + 2 * d) : (long)d * (g.nx + 2 * d);
    long end = 0;
    int n = 0;
    for (int f = 0; f < TEA_NUM_FIELDS; f++) {
        if (!fields[f]) continue;
        if (offsets[f] < 0 || offsets[f] > TEA_NUM_FIELDS * side) {
            die("Tried to pack/unpack field %d but invalid offset %d given", f, offsets[f]);
            return;
        }
        if (offsets[f] + side > end) end = offsets[f] + side;
        n++;
    }
    if (n == 0) return;
    if ((size_t)end > g.halo_buf_elems) { die("pack: buffer of %ld elements exceeds staging", end); return; }
    if (!pack && fields[TEA_FIELD_SD - 1]) g.sd_halo_dirty = true;
    if (!pack && fields[TEA_FIELD_P - 1]) g.p_halo_dirty[0] = true;
    if (!pack) HIP_OK(hipMemcpyAsync(g.d_halo_buf, host_buffer, end * sizeof(double), hipMemcpyHostToDevice, g.stream));
    const int len = lr ? g.ny + 2 * d : g.nx + 2 * d;
    for (int f = 0; f < TEA_NUM_FIELDS; f++) {
        if (!fields[f]) continue;
        Timed t(pack ? "pack" : "unpack");
        hipLaunchKernelGGL(k_pack, dim3((len + 255) / 256, d), dim3(256), 0, g.stream, F(f + 1),
                           g.d_halo_buf + offsets[f], g.pitch, g.h, g.nx, g.ny, d, face, pack);
    }
    HIP_OK(hipGetLastError());
    if (pack) HIP_OK(hipMemcpyAsync(host_buffer, g.d_halo_buf, end * sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIP_OK(hipStreamSynchronize(g.stream));
}
}  // namespace

EXPORT void cuda_pack_buffers_(int *fields, int *offsets, int *depth, int *face, double *host_buffer)
{
    if (g.initialised) settle_pending();   // rtemp and utemp owed by tea_engine_ppcg_begin_ur
    pack_unpack(fields, offsets, *depth, *face, 1, host_buffer);
}

EXPORT void cuda_unpack_buffers_(int *fields, int *offsets, int *depth, int *face, double *host_buffer)
{
    if (g.initialised) settle_pending();   // rtemp and utemp owed by tea_engine_ppcg_begin_ur
    g.u_ring_dirty = true;                 // halo cells of u may change
    pack_unpack(fields, offsets, *depth, *face, 0, host_buffer);
}

EXPORT void tea_leaf_common_init_kernel_cuda_(const int *coefficient, double *dt, double *rx, double *ry,
                                              int *zero_boundary, int *reflective_boundary)
{
    CHECK_INIT();
    const int coef = *coefficient;
    if (coef != TEA_CONDUCTIVITY && coef != TEA_RECIP_CONDUCTIVITY) {
        die("Unknown coefficient %d passed to tea leaf", coef);
        return;
    }
    // calcrxry, tea_leaf_kernel_cuda.cu:16-32 (dx, dy kept on the host)
    const double dxv = g.h_celldx[g.h], dyv = g.h_celldy[g.h];
    *rx = *dt / (dxv * dxv);
    *ry = *dt / (dyv * dyv);
    const int h = g.h;
    // index of cell (0,0) relative to the shifted base, for the reference's
    // "THARR2D(-1,0,0) > 0" guards
    const long i_origin = 0;
    launch("init_common", k_init_common, Ext{h, h + g.nx - 1 + h, h, h + g.ny - 1 + h}, g.pitch,
           (const double *)F(F_DENSITY), F(X_KX), F(X_KY), *rx, *ry, coef, i_origin);
    if (!*reflective_boundary) {
        launch("zero_boundaries", k_zero_boundaries, Ext{0, g.nx + 2 * h - 1, 0, g.ny + 2 * h - 1}, g.pitch,
               F(X_KX), F(X_KY), h, g.nx, g.ny, zero_boundary[0], zero_boundary[1], zero_boundary[2],
               zero_boundary[3]);
    }
    launch("generate_chunk_init_u", k_init_u, interior(), g.pitch, (const double *)F(F_DENSITY),
           (const double *)F(F_ENERGY1), F(F_U), F(X_U0));
}

EXPORT void tea_leaf_calc_residual_cuda_(void)
{
    CHECK_INIT();
    launch("calc_residual", k_calc_residual, interior(), g.pitch, (const double *)F(F_U),
           (const double *)F(X_U0), F(F_R), (const double *)F(X_KX), (const double *)F(X_KY));
}

EXPORT void tea_leaf_common_finalise_kernel_cuda_(void)
{
    CHECK_INIT();
    launch("finalise", k_finalise, interior(), g.pitch, (const double *)F(F_DENSITY), (const double *)F(F_U),
           F(F_ENERGY1));
}

EXPORT void tea_leaf_calc_2norm_kernel_cuda_(int *norm_array, double *norm)
{
    CHECK_INIT_KEEP_PENDING();
    g.u_ring_dirty = true;
    if (*norm_array < 0 || *norm_array > 2) {
        die("Invalid value '%d' for norm_array passed, should be [0, 1, 2]", *norm_array);
        return;
    }
    if (*norm_array == 1 && g.pending_cheby == (g.fp32 ? 32 : 64)) {
        // r.r of the r a fused Chebyshev iteration owes, formed on the fly (it stays owed)
        bool ok;
        if (g.fp32)
            ok = launch("calc_2norm_cheby_f32", k_cheby_rr<float>, interior(), g.pitch32,
                        (const float *)g.fs32[X_U2], (const float *)g.fs32[X_U0], (const float *)g.fs32[X_KX],
                        (const float *)g.fs32[X_KY], red_ctx());
        else
            ok = launch("calc_2norm_cheby", k_cheby_rr<double>, interior(), g.pitch, (const double *)F(X_U2),
                        (const double *)F(X_U0), (const double *)F(X_KX), (const double *)F(X_KY), red_ctx());
        if (ok) fetch(norm, 1);
        return;
    }
    settle_pending();
    const int u0 = X_U0, r = F_R, z = g.prec != PREC_NONE ? X_Z : F_R;
    const int ia = *norm_array == 0 ? u0 : r, ib = *norm_array == 0 ? u0 : (*norm_array == 1 ? r : z);
    bool ok;
    if (g.fp32)   // the FP32 solver's fields (its right-hand side, residual, z)
        ok = launch("calc_2norm_f32", k_dot<float>, interior(), g.pitch32, (const float *)g.fs32[ia],
                    (const float *)g.fs32[ib], red_ctx());
    else
        ok = launch("calc_2norm", k_dot<double>, interior(), g.pitch, (const double *)F(ia), (const double *)F(ib),
                    red_ctx());
    if (ok) fetch(norm, 1);
}

// ---- the 2-norm without allocation ---------------------------------------------
//
// tea_engine_calc_2norm_noalloc: the dot product of tea_leaf_calc_2norm on
// caller-owned device arrays, for callers that cannot allocate (code under
// graph capture, serving loops).  Its workgroup partials live in `scratch`, a
// file-scope device array of NOALLOC_BLOCKS entries (64 KiB) that each GPU
// holds until the process ends; the grid is capped at that many workgroups,
// which then loop over tiles, so any mesh size fits.  Every call on one GPU
// shares that GPU's array: calls that overlap on one GPU race.
namespace {
constexpr int NOALLOC_BLOCKS = 4096;
constexpr int NOALLOC_MAX_DEVICES = 64;
constexpr int NOALLOC_MAX_CELLS_SIDE = 1 << 16;
static __device__ __align__(256) DD scratch[NOALLOC_BLOCKS];
// The device address of each GPU's copy of `scratch` (the host-side name of a
// __device__ array is no device address), looked up once per device.
DD *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {nullptr};

DD *scratch_on_current_device(char *err, size_t errlen)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        snprintf(err, errlen, "hipGetDevice failed: %s", hipGetErrorString(e));
        return nullptr;
    }
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        snprintf(err, errlen, "device %d beyond the %d devices supported", dev, NOALLOC_MAX_DEVICES);
        return nullptr;
    }
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            snprintf(err, errlen, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = (DD *)p;
    }
    return g_scratch_ptr[dev];
}

// Errors of the no-alloc entry point are returned, never fatal: the message
// goes to stderr and to the sticky error (tea_engine_last_error).
int noalloc_error(const char *msg)
{
    snprintf(g_error, sizeof g_error, "calc_2norm_noalloc: %s", msg);
    fprintf(stderr, "tea_engine: %s\n", g_error);
    fflush(stderr);
    return -1;
}
}  // namespace

EXPORT int tea_engine_calc_2norm_noalloc(const double *a, const double *b, int nx, int ny, int halo, long pitch,
                                         double *dev_norm, void *stream)
{
    char err[256];
    if (!a || !b || !dev_norm) return noalloc_error("null array or result pointer");
    if (nx < 1 || ny < 1 || nx > NOALLOC_MAX_CELLS_SIDE || ny > NOALLOC_MAX_CELLS_SIDE || halo < 0 ||
        halo > 64 || pitch < (long)nx + 2L * halo) {
        snprintf(err, sizeof err, "mesh %d x %d, halo %d, pitch %ld: needs 1..%d cells per side, halo 0..64, "
                 "pitch >= nx + 2 halo", nx, ny, halo, pitch, NOALLOC_MAX_CELLS_SIDE);
        return noalloc_error(err);
    }
    DD *partials = scratch_on_current_device(err, sizeof err);
    if (!partials) return noalloc_error(err);
    const Ext e{halo, halo + nx - 1, halo, halo + ny - 1};
    const int tiles = blocks_for(e);
    const int grid = tiles < NOALLOC_BLOCKS ? tiles : NOALLOC_BLOCKS;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_dot<double>, dim3(grid), dim3(BLOCK), 0, s, e, pitch, a, b, Red{partials, NOALLOC_BLOCKS});
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(FINISH_THREADS), 0, s, (const DD *)partials, NOALLOC_BLOCKS, grid, 1,
                       dev_norm, (double *)nullptr);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
        snprintf(err, sizeof err, "launch failed: %s", hipGetErrorString(le));
        return noalloc_error(err);
    }
    return 0;
}

namespace {

// Fields in the precision of the current solver mode: FP64 (the reference's
// path) or FP32 (the inner solver of the mixed-precision refinement).
template <typename T> T *FT(int id);
template <> double *FT<double>(int id) { return g.fs[id]; }
template <> float *FT<float>(int id) { return g.fs32[id]; }
template <typename T> long pitch_of() { return std::is_same<T, float>::value ? g.pitch32 : g.pitch; }

// Profiler name of a launch: FP32 launches carry a "_f32" suffix.
template <typename T> const char *nm(const char *base)
{
    if (!std::is_same<T, float>::value) return base;
    static std::map<std::string, std::string> names;
    auto it = names.find(base);
    if (it == names.end()) it = names.emplace(base, std::string(base) + "_f32").first;
    return it->second.c_str();
}

// Writes out rtemp = r and utemp = sd owed by tea_engine_ppcg_begin_ur (on the
// interior, as ppcg_init_sd_new writes them).
template <typename T> void settle_fields()
{
    const Ext in = interior();
    launch(nm<T>("settle_rtemp"), k_copy<T>, in, pitch_of<T>(), FT<T>(X_RTEMP), (const T *)FT<T>(F_R));
    launch(nm<T>("settle_utemp"), k_copy<T>, in, pitch_of<T>(), FT<T>(X_UTEMP), (const T *)FT<T>(F_SD));
}

// z = M^-1 r (z) and sd = z / theta (sd) owed by a fused ur, written out
void settle_minv_r(bool z, bool sd)
{
    const int zb = z ? g.pending_zr : 0, sb = sd ? g.pending_sd_r : 0;
    if (z) g.pending_zr = 0;
    if (sd) g.pending_sd_r = 0;
    const int bits = zb ? zb : sb;
    if (g.prec == PREC_DIAG) {   // z owed by a fused CG iteration (FP64; no sd)
        if (zb == 64)
            launch("settle_minv_r_diag", k_minv_r_diag<double>, interior(), g.pitch, (const double *)g.fs[F_R],
                   g.fs[X_Z], (const double *)g.fs[X_KX], (const double *)g.fs[X_KY]);
        return;
    }
    if (bits == 32)
        launch("settle_minv_r_f32", k_minv_r<float>, interior(), g.pitch32, (const float *)g.fs32[F_R],
               zb ? g.fs32[X_Z] : (float *)nullptr, sb ? g.fs32[F_SD] : (float *)nullptr, g.sd_r_theta,
               (const float *)g.fs32[X_KX], (const float *)g.fs32[X_KY]);
    else if (bits == 64)
        launch("settle_minv_r", k_minv_r<double>, interior(), g.pitch, (const double *)g.fs[F_R],
               zb ? g.fs[X_Z] : (double *)nullptr, sb ? g.fs[F_SD] : (double *)nullptr, g.sd_r_theta,
               (const double *)g.fs[X_KX], (const double *)g.fs[X_KY]);
}

// z = M^-1 r owed (only)
void settle_z_r() { settle_minv_r(true, false); }

// z = utemp owed (only): a z = M^-1 r debt stays owed
void settle_z_utemp()
{
    const int bits = g.pending_z;
    g.pending_z = 0;
    if (bits == 32)
        launch("settle_z_f32", k_copy<float>, interior(), g.pitch32, g.fs32[X_Z], (const float *)g.fs32[X_UTEMP]);
    else if (bits == 64)
        launch("settle_z", k_copy<double>, interior(), g.pitch, g.fs[X_Z], (const double *)g.fs[X_UTEMP]);
}

// z owed (only): z = utemp, or z = M^-1 r
void settle_z()
{
    settle_minv_r(true, false);
    settle_z_utemp();
}

// rtemp and utemp owed (only; an sd they copy that is itself owed first)
void settle_rt_ut()
{
    if (g.pending_rt_ut) settle_minv_r(false, true);
    const int bits = g.pending_rt_ut;
    g.pending_rt_ut = 0;
    if (bits == 32) settle_fields<float>();
    else if (bits == 64) settle_fields<double>();
}

void settle_pending()
{
    settle_minv_r(true, true);   // one launch for z and sd owed together
    settle_z();   // before the rtemp/utemp debt rewrites utemp
    settle_rt_ut();
    const int cbits = g.pending_cheby;
    g.pending_cheby = 0;
    if (cbits == 32)
        launch("settle_cheby_wr_f32", k_cheby_wr<float>, interior(), g.pitch32, (const float *)g.fs32[X_U2],
               (const float *)g.fs32[X_U0], g.fs32[F_R], g.fs32[X_W], (const float *)g.fs32[X_KX],
               (const float *)g.fs32[X_KY]);
    else if (cbits == 64)
        launch("settle_cheby_wr", k_cheby_wr<double>, interior(), g.pitch, (const double *)g.fs[X_U2],
               (const double *)g.fs[X_U0], g.fs[F_R], g.fs[X_W], (const double *)g.fs[X_KX],
               (const double *)g.fs[X_KY]);
    const int wbits = g.pending_w;   // w = A p as cg_calc_w forms it (its p.w is not used)
    g.pending_w = 0;
    if (wbits == 32)
        launch("settle_w_f32", k_cg_calc_w<float>, interior(), g.pitch32, (const float *)g.fs32[F_P], g.fs32[X_W],
               (const float *)g.fs32[X_KX], (const float *)g.fs32[X_KY], red_ctx());
    else if (wbits == 64)
        launch("settle_w", k_cg_calc_w<double>, interior(), g.pitch, (const double *)g.fs[F_P], g.fs[X_W],
               (const double *)g.fs[X_KX], (const double *)g.fs[X_KY], red_ctx());
}

#define PREC_SWITCH_T(T, PREC_VALUE, NAME, KERNEL, ...)                                           \
    ((PREC_VALUE) == PREC_BLOCK ? launch(nm<T>(NAME), KERNEL<T, PREC_BLOCK>, __VA_ARGS__)          \
     : (PREC_VALUE) == PREC_DIAG ? launch(nm<T>(NAME), KERNEL<T, PREC_DIAG>, __VA_ARGS__)          \
                                 : launch(nm<T>(NAME), KERNEL<T, PREC_NONE>, __VA_ARGS__))

template <typename T> void cg_init_impl(double *rro)
{
    const Ext in = interior();
    const long pt = pitch_of<T>();
    if (g.prec == PREC_BLOCK && std::is_same<T, double>::value)
        launch("block_init", k_block_init<T>, in, pt, FT<T>(X_CP), FT<T>(X_BFP), (const T *)FT<T>(X_KX),
               (const T *)FT<T>(X_KY));
    else if (g.prec == PREC_DIAG)
        launch(nm<T>("init_jac_diag"), k_jac_diag<T>, in, pt, FT<T>(X_MI), (const T *)FT<T>(X_KX),
               (const T *)FT<T>(X_KY));
    if (!PREC_SWITCH_T(T, g.prec, "k_cg_init_p", k_cg_init_p, in, pt, FT<T>(F_P), (const T *)FT<T>(F_R), FT<T>(X_Z),
                       (const T *)FT<T>(X_MI), (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), red_ctx()))
        return;
    fetch(rro, 1);
}

template <typename T> void cg_calc_w_impl(double *pw)
{
    if (!launch(nm<T>("cg_calc_w"), k_cg_calc_w<T>, interior(), pitch_of<T>(), (const T *)FT<T>(F_P), FT<T>(X_W),
                (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), red_ctx()))
        return;
    fetch(pw, 1);
}

template <typename T> void cg_calc_ur_impl(double alpha, double *rrn)
{
    if (!PREC_SWITCH_T(T, g.prec, "k_cg_calc_ur", k_cg_calc_ur, interior(), pitch_of<T>(), alpha, FT<T>(F_U),
                       (const T *)FT<T>(F_P), FT<T>(F_R), (const T *)FT<T>(X_W), FT<T>(X_Z), (const T *)FT<T>(X_MI),
                       (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), red_ctx()))
        return;
    fetch(rrn, 1);
}

template <typename T> void cg_calc_p_impl(double beta)
{
    launch(nm<T>("cg_calc_p"), k_calc_p<T>, interior(), pitch_of<T>(), beta, FT<T>(F_P),
           (const T *)(g.prec != PREC_NONE ? FT<T>(X_Z) : FT<T>(F_R)));
}

template <typename T> void ppcg_init_impl(int st, double *rro)
{
    if (!PREC_SWITCH_T(T, g.prec, "k_ppcg_init", k_ppcg_init, interior(), pitch_of<T>(), FT<T>(F_P),
                       (const T *)FT<T>(F_R), FT<T>(X_Z), (const T *)FT<T>(X_MI), (const T *)FT<T>(X_KX),
                       (const T *)FT<T>(X_KY), st, red_ctx()))
        return;
    double v = 0.0;
    if (!fetch(&v, 1)) return;
    *rro = (st == 2) ? 0.0 : v;
}

template <typename T> void ppcg_init_sd_impl(double theta)
{
    PREC_SWITCH_T(T, g.prec, "k_ppcg_init_sd", k_ppcg_init_sd, interior(), pitch_of<T>(), (const T *)FT<T>(F_R),
                  FT<T>(F_SD), (const T *)FT<T>(X_MI), (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), theta);
}

template <typename T> void ppcg_init_sd_new_impl(double theta)
{
    launch(nm<T>("ppcg_init_sd_new"), k_ppcg_init_sd_new<T>, interior(), pitch_of<T>(), (const T *)FT<T>(F_R),
           FT<T>(F_SD), (const T *)(g.prec != PREC_NONE ? FT<T>(X_Z) : FT<T>(F_R)), FT<T>(X_RTEMP),
           FT<T>(X_UTEMP), theta);
}

template <typename T> void ppcg_inner_impl(const Ext &e, double alpha, double beta)
{
    const long pt = pitch_of<T>();
    if (fused_inner()) {
        bool &dirty = std::is_same<T, double>::value ? g.sd_halo_dirty : g.sd32_halo_dirty;
        if (dirty) {
            const long ring = 2L * (g.nx + 2L * g.h) * g.h + 2L * g.ny * g.h;
            Timed t("sd_halo_ring");
            hipLaunchKernelGGL(k_copy_ring<T>, dim3((unsigned)((ring + 255) / 256)), dim3(256), 0, g.stream,
                               FT<T>(X_SD2), (const T *)FT<T>(F_SD), pt, g.h, g.nx, g.ny);
            dirty = false;
        }
        // the new sd goes to the spare buffer, which then becomes the sd field
        if (!PREC_SWITCH_T(T, g.prec, "k_ppcg_inner_fused", k_ppcg_inner_fused, e, pt, (const T *)FT<T>(F_SD),
                           FT<T>(X_SD2), FT<T>(X_RTEMP), FT<T>(X_UTEMP), (const T *)FT<T>(X_MI),
                           (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), alpha, beta))
            return;
        if (std::is_same<T, double>::value) {
            std::swap(g.f[F_SD], g.f[X_SD2]);
            std::swap(g.fs[F_SD], g.fs[X_SD2]);
        } else {
            std::swap(g.fs32[F_SD], g.fs32[X_SD2]);
        }
        return;
    }
    if (!launch(nm<T>("ppcg_update_r"), k_ppcg_update_r<T>, e, pt, FT<T>(X_RTEMP), (const T *)FT<T>(X_KX),
                (const T *)FT<T>(X_KY), (const T *)FT<T>(F_SD)))
        return;
    PREC_SWITCH_T(T, g.prec, "k_ppcg_calc_sd", k_ppcg_calc_sd, e, pt, FT<T>(F_SD), (const T *)FT<T>(X_RTEMP),
                  FT<T>(X_UTEMP), (const T *)FT<T>(X_MI), (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), alpha,
                  beta);
}

// NS inner steps in one pass (k_ppcg_inner_n_x2 for FP32, k_ppcg_inner_n for
// FP64) on the interior.  16 waves per workgroup measured fastest on MI355X
// for every NS and both field types (against 8, 10 and 12).
constexpr int INNER_WAVES = 16;

// Microseconds per pass of NS = 1..MAX_PASS_STEPS inner steps at 4096^2 on
// MI355X (back-to-back passes over the same fields), with the kernel forms
// inner_form picks: the plan of passes for a run of inner steps minimises
// their sum.
constexpr double PASS_US_F32[MAX_PASS_STEPS] = {128, 138, 146, 150, 185};
constexpr double PASS_US_F64[MAX_PASS_STEPS] = {239, 243, 255, 306, 766};
constexpr double PASS_US_F32_DIAG[MAX_PASS_STEPS] = {144, 152, 161, 161, 168};

const double *pass_costs()
{
    if (!g.fp32) return PASS_US_F64;
    return g.prec == PREC_DIAG ? PASS_US_F32_DIAG : PASS_US_F32;
}

// Which form of the streaming pass runs (all give the same results): 0 the
// shifting pipeline (k_ppcg_inner_march), 1 the ring (k_ppcg_inner_ring),
// 2 the ring with the coefficients in LDS (k_ppcg_inner_ringl) -- the
// fastest measured at 4096^2 for each field type, preconditioner and number
// of steps; TEA_INNER_RING=0|1|2 forces one.
template <typename T, int PREC, int NS> constexpr int inner_form()
{
    if (sizeof(T) == 4) return NS == 5 ? 2 : 1;
    if (NS <= 2) return 0;
    if (NS == 3) return 1;
    return PREC == PREC_BLOCK ? 2 : 1;
}

// The streaming pass (k_ppcg_inner_march) is the faster kernel for every NS
// in both precisions (measured against k_ppcg_inner_n[_x2]); it addresses
// cells with 32-bit byte offsets, so fields must be under 4 GiB.  Five FP64
// steps per pass, whose pipeline spills most of its registers into AGPRs,
// run on the tile kernel: with separate utemp read and write resources and
// utemp read only for the wave's own strips that variant gave results that
// changed from run to run, and no plan uses it (ten FP64 steps go as 4+3+3).
// TEA_INNER_MARCH=0 selects the tile kernels.
template <typename T, int NS> bool use_march()
{
    const char *s = getenv("TEA_INNER_MARCH");
    if (s && s[0] == '0') return false;
    if (sizeof(T) == 8 && NS == 5) return false;
    return (double)pitch_of<T>() * (double)g.rows * sizeof(T) < 4294967296.0 - 65536.0;
}

// Waves per SIMD the row-streaming kernels are sized for (longer runs re-read
// fewer rows, shorter ones hide more latency): the measured best at 4096^2,
// 8 for k_begin_march and 4 for k_cheby_march (199 against 212 us per
// Chebyshev iteration at 8), or the value of the named environment variable.
int stream_waves(const char *env, int def)
{
    const char *s = getenv(env);
    const int v = s ? atoi(s) : 0;
    return v > 0 && v <= 32 ? v : def;
}

int cu_count()
{
    static int v = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1)
            n = 256;
        return n;
    }();
    return v;
}

// The passes that can form sd = M^-1 r / theta from r themselves
// (k_ppcg_inner_ringl FROM_R): the first pass of ten block-Jacobi inner steps
// (5 + 5 in FP32, 4 + 3 + 3 in FP64).
template <typename T, int NS> constexpr bool has_from_r() { return sizeof(T) == 4 ? NS == 5 : NS == 4; }

// One wave per band of columns and chunk of strips: two waves per SIMD in
// FP32 (its kernels fit 256 registers up to five steps), one in FP64 (the
// longer chunks that leaves measured faster than two waves per SIMD even
// where the registers would allow two), chunks as long as that leaves.
// With r_end (and rs_end) given and a ring form with a preconditioner
// running, the pass also forms k_ppcg_end's (r - rstore).z over the utemp it
// writes (z owed, not stored) into the reduction partials, and returns true.
// sd owed by a fused ur (with rtemp and utemp): the LDS ring pass forms it
// from r where it can (FROM_R), else it is written out first.
template <typename T, int NS>
bool launch_inner_march(const Ext &e, long pt, const StepCoefs &cf_in, const char *name, const T *r_end = nullptr,
                        const T *rs_end = nullptr)
{
    const char *ring_env = getenv("TEA_INNER_RING");
    const int forced = ring_env && ring_env[0] >= '0' && ring_env[0] <= '2' ? ring_env[0] - '0' : -1;
    const int bits = 8 * (int)sizeof(T);
    bool from_r = false;
    if (g.pending_sd_r == bits) {
        from_r = has_from_r<T, NS>() && g.prec == PREC_BLOCK && g.pending_rt_ut == bits && !(r_end && rs_end) &&
                 (forced >= 0 ? forced : inner_form<T, PREC_BLOCK, NS>()) == 2;
        if (!from_r) settle_minv_r(false, true);
        g.pending_sd_r = 0;
    }
    StepCoefs cf = cf_in;
    cf.theta = g.sd_r_theta;
    const int OX = march_out_x(NS, from_r);
    const int nbands = (e.x1 - e.x0 + OX) / OX, nstrips = (e.y1 - e.y0 + STRIP) / STRIP;
    if (nbands < 1 || nstrips < 1) return false;
    const int per_simd = sizeof(T) == 4 ? 2 : 1;
    int nchunks = cu_count() * 4 * per_simd / nbands;
    nchunks = nchunks < 1 ? 1 : (nchunks > nstrips ? nstrips : nchunks);
    int chunk = (nstrips + nchunks - 1) / nchunks;
    // TEA_MARCH_CHUNK=<strips> sets the chunk length (tests run the
    // production shape's multi-strip chunks on small meshes with it)
    if (const char *s = getenv("TEA_MARCH_CHUNK")) {
        const int c = atoi(s);
        if (c > 0) chunk = c < nstrips ? c : nstrips;
    }
    nchunks = (nstrips + chunk - 1) / chunk;
    const unsigned grid = (unsigned)((nbands * nchunks + MARCH_WAVES - 1) / MARCH_WAVES);
    // rtemp = r and utemp = sd still owed by tea_engine_ppcg_begin_ur: read r and sd instead
    const bool owed = g.pending_rt_ut == 8 * (int)sizeof(T);
    const T *rt_in = owed ? FT<T>(F_R) : FT<T>(X_RTEMP), *ut_in = owed ? FT<T>(F_SD) : FT<T>(X_UTEMP);
    g.pending_rt_ut = 0;
    auto go = [&](auto kernel) {
        klaunch(name, kernel, dim3(grid), dim3(MARCH_WAVES * WAVE), e, pt, chunk, (const T *)FT<T>(F_SD),
                FT<T>(X_SD2), rt_in, FT<T>(X_RT2), ut_in, FT<T>(X_UTEMP), (const T *)FT<T>(X_MI),
                (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), cf);
    };
    // the instances with the end's reduction are timed under their own name
    static const char *end_names[MAX_PASS_STEPS] = {"k_ppcg_inner1_end", "k_ppcg_inner2_end", "k_ppcg_inner3_end",
                                                    "k_ppcg_inner4_end", "k_ppcg_inner5_end"};
    static const char *from_r_names[MAX_PASS_STEPS] = {"k_ppcg_inner1_from_r", "k_ppcg_inner2_from_r",
                                                       "k_ppcg_inner3_from_r", "k_ppcg_inner4_from_r",
                                                       "k_ppcg_inner5_from_r"};
    auto go_ring = [&](auto kernel, const T *re, const T *rse) {
        g.last_grid = (int)grid;
        const char *nmr = re ? nm<T>(end_names[NS - 1]) : from_r ? nm<T>(from_r_names[NS - 1]) : name;
        klaunch(nmr, kernel, dim3(grid), dim3(MARCH_WAVES * WAVE), e, pt, chunk, (const T *)FT<T>(F_SD),
                FT<T>(X_SD2), rt_in, FT<T>(X_RT2), ut_in, FT<T>(X_UTEMP), (const T *)FT<T>(X_MI),
                (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), cf, re, rse, red_ctx());
    };
    const bool want_end = r_end && rs_end && (int)grid <= g.max_blocks;
    bool end_done = false;
    auto pick = [&](auto prec_c) {
        constexpr int P = decltype(prec_c)::value;
        const int form = forced >= 0 ? forced : inner_form<T, P, NS>();
        if constexpr (P == PREC_BLOCK && has_from_r<T, NS>()) {
            if (from_r) {   // chosen above: form 2, no end
                go_ring(k_ppcg_inner_ringl<T, P, NS, false, true>, nullptr, nullptr);
                return;
            }
        }
        if constexpr (P != PREC_NONE) {
            if (want_end && form == 2) {
                go_ring(k_ppcg_inner_ringl<T, P, NS, true>, r_end, rs_end);
                end_done = true;
                return;
            }
            if (want_end && form == 1) {
                go_ring(k_ppcg_inner_ring<T, P, NS, true>, r_end, rs_end);
                end_done = true;
                return;
            }
        }
        if (form == 2) go_ring(k_ppcg_inner_ringl<T, P, NS>, nullptr, nullptr);
        else if (form == 1) go_ring(k_ppcg_inner_ring<T, P, NS>, nullptr, nullptr);
        else go(k_ppcg_inner_march<T, P, NS>);
    };
    if (g.prec == PREC_BLOCK) pick(std::integral_constant<int, PREC_BLOCK>{});
    else if (g.prec == PREC_DIAG) pick(std::integral_constant<int, PREC_DIAG>{});
    else pick(std::integral_constant<int, PREC_NONE>{});
    return end_done;
}

int plan_first_pass(int remaining, const double *cost)
{
    if (remaining <= 0) return 0;
    std::vector<double> best((size_t)remaining + 1, 0.0);
    std::vector<int> first((size_t)remaining + 1, 0);
    for (int r = 1; r <= remaining; r++) {
        best[r] = 1e300;
        for (int n = std::min(r, MAX_PASS_STEPS); n >= 1; n--) {   // ties: the longer pass first
            const double v = cost[n - 1] + best[r - n];
            if (v < best[r]) {
                best[r] = v;
                first[r] = n;
            }
        }
    }
    return first[remaining];
}

template <typename T, int NS>
bool launch_inner_n(const Ext &e, long pt, const StepCoefs &cf, const T *r_end, const T *rs_end)
{
    constexpr bool X2 = std::is_same<T, float>::value;
    using Tile = InnerTile<X2, INNER_WAVES, NS>;
    static const char *names[MAX_PASS_STEPS] = {"k_ppcg_inner1", "k_ppcg_inner2", "k_ppcg_inner3", "k_ppcg_inner4",
                                                "k_ppcg_inner5"};
    if (use_march<T, NS>()) return launch_inner_march<T, NS>(e, pt, cf, nm<T>(names[NS - 1]), r_end, rs_end);
    settle_pending();
    const long ntiles = (long)((e.x1 - e.x0 + Tile::OUT_X) / Tile::OUT_X) * ((e.y1 - e.y0 + Tile::OUT_Y) / Tile::OUT_Y);
    const unsigned grid = (unsigned)std::min<long>(ntiles, grid_max());
    auto go = [&](auto kernel) {
        klaunch(nm<T>(names[NS - 1]), kernel, dim3(grid), dim3(INNER_WAVES * WAVE), e, pt, (const T *)FT<T>(F_SD),
                FT<T>(X_SD2), (const T *)FT<T>(X_RTEMP), FT<T>(X_RT2), FT<T>(X_UTEMP), (const T *)FT<T>(X_MI),
                (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), cf);
    };
    if constexpr (X2) {
        if (g.prec == PREC_BLOCK) go(k_ppcg_inner_n_x2<PREC_BLOCK, INNER_WAVES, NS>);
        else if (g.prec == PREC_DIAG) go(k_ppcg_inner_n_x2<PREC_DIAG, INNER_WAVES, NS>);
        else go(k_ppcg_inner_n_x2<PREC_NONE, INNER_WAVES, NS>);
    } else {
        if (g.prec == PREC_BLOCK) go(k_ppcg_inner_n<T, PREC_BLOCK, INNER_WAVES, NS>);
        else if (g.prec == PREC_DIAG) go(k_ppcg_inner_n<T, PREC_DIAG, INNER_WAVES, NS>);
        else go(k_ppcg_inner_n<T, PREC_NONE, INNER_WAVES, NS>);
    }
    return false;
}

// Steps k0+1 .. k0+ns (0-based coefficient index k0) in one pass.  with_end:
// the pass may also form the end's (r - rstore).z (returns whether it did; its
// partials are then the latest reduction's).
template <typename T> bool ppcg_inner_n_impl(int k0, int ns, bool with_end = false)
{
    const long pt = pitch_of<T>();
    const long ring = 2L * (g.nx + 2L * g.h) * g.h + 2L * g.ny * g.h;
    const unsigned ring_blocks = (unsigned)((ring + 255) / 256);
    bool &dirty = std::is_same<T, double>::value ? g.sd_halo_dirty : g.sd32_halo_dirty;
    bool &rt_dirty = g.rt_halo_dirty[std::is_same<T, double>::value ? 0 : 1];
    if (dirty || rt_dirty) {
        Timed t("inner_halo_rings");
        if (dirty) {
            hipLaunchKernelGGL(k_copy_ring<T>, dim3(ring_blocks), dim3(256), 0, g.stream, FT<T>(X_SD2),
                               (const T *)FT<T>(F_SD), pt, g.h, g.nx, g.ny);
            dirty = false;
        }
        if (rt_dirty) {
            hipLaunchKernelGGL(k_copy_ring<T>, dim3(ring_blocks), dim3(256), 0, g.stream, FT<T>(X_RT2),
                               (const T *)FT<T>(X_RTEMP), pt, g.h, g.nx, g.ny);
            rt_dirty = false;
        }
    }
    StepCoefs cf{};
    for (int k = 0; k < ns; k++) {
        cf.a[k] = g.ch_a[k0 + k];
        cf.b[k] = g.ch_b[k0 + k];
    }
    const Ext e = interior();
    const T *re = with_end ? (const T *)FT<T>(F_R) : nullptr, *rse = with_end ? (const T *)FT<T>(X_RSTORE) : nullptr;
    bool end_done = false;
    switch (ns) {
    case 1: end_done = launch_inner_n<T, 1>(e, pt, cf, re, rse); break;
    case 2: end_done = launch_inner_n<T, 2>(e, pt, cf, re, rse); break;
    case 3: end_done = launch_inner_n<T, 3>(e, pt, cf, re, rse); break;
    case 4: end_done = launch_inner_n<T, 4>(e, pt, cf, re, rse); break;
    default: end_done = launch_inner_n<T, 5>(e, pt, cf, re, rse); break;
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) { die("launch of k_ppcg_inner%d failed: %s", ns, hipGetErrorString(err)); return false; }
    if (std::is_same<T, double>::value) {
        std::swap(g.f[F_SD], g.f[X_SD2]);
        std::swap(g.fs[F_SD], g.fs[X_SD2]);
        std::swap(g.f[X_RTEMP], g.f[X_RT2]);
        std::swap(g.fs[X_RTEMP], g.fs[X_RT2]);
    } else {
        std::swap(g.fs32[F_SD], g.fs32[X_SD2]);
        std::swap(g.fs32[X_RTEMP], g.fs32[X_RT2]);
    }
    return end_done;
}

template <typename T> void ppcg_store_r_impl()
{
    launch(nm<T>("ppcg_store_r"), k_copy<T>, interior(), pitch_of<T>(), FT<T>(X_RSTORE), (const T *)FT<T>(F_R));
}

template <typename T> void ppcg_update_z_impl()
{
    launch(nm<T>("ppcg_update_z"), k_copy<T>, interior(), pitch_of<T>(), FT<T>(X_Z), (const T *)FT<T>(X_UTEMP));
}

template <typename T> void ppcg_calc_rrn_impl(double *norm)
{
    if (!launch(nm<T>("ppcg_calc_rrn"), k_calc_rrn<T>, interior(), pitch_of<T>(), (const T *)FT<T>(X_RSTORE),
                (const T *)FT<T>(F_R), (const T *)FT<T>(X_Z), red_ctx()))
        return;
    fetch(norm, 1);
}

template <typename T> void ppcg_calc_2norm_impl(double *norm)
{
    if (!launch(nm<T>("ppcg_calc_2norm"), k_dot<T>, interior(), pitch_of<T>(), (const T *)FT<T>(F_R),
                (const T *)FT<T>(X_Z), red_ctx()))
        return;
    fetch(norm, 1);
}

template <typename T> void ppcg_calc_p_impl(double beta)
{
    launch(nm<T>("ppcg_calc_p"), k_calc_p<T>, interior(), pitch_of<T>(), beta, FT<T>(F_P), (const T *)FT<T>(X_Z));
}

// Fused outer iteration (see k_ppcg_begin / k_ppcg_ur_sd / k_ppcg_end).
// beta_num: null (beta given), or device memory whose value over beta is the
// beta to apply (formed on the device as the host would form it).
// Block-Jacobi fused pair: begin leaves w = A p' unwritten and ur forms it
// again from p' (k_ppcg_ur_sd_x2_wp, k_ppcg_ur_sd_wp), unless TEA_W_FROM_P=0.
template <typename T> bool w_from_p()
{
    const char *s = getenv("TEA_W_FROM_P");
    return g.prec == PREC_BLOCK && !(s && s[0] == '0');
}

// Whether a fused ur leaves z and sd to the first inner pass (block Jacobi,
// w formed from p, rtemp and utemp deferred to a streaming pass) unless
// TEA_SD_FROM_R=0: the pass forms sd from r if its form can (else the
// values are written out first).
template <typename T> bool sd_from_r()
{
    const char *s = getenv("TEA_SD_FROM_R");
    return g.prec == PREC_BLOCK && w_from_p<T>() && use_march<T, 1>() && !(s && s[0] == '0');
}

// The streaming form of the fused begin (k_begin_march) unless
// TEA_BEGIN_MARCH=0: bands of 62 columns, runs of rows sized for about eight
// waves per SIMD.  Returns false (nothing launched) when it does not apply.
template <typename T, bool STORE_R, bool Z_FROM_R = false>
bool begin_march_launch(const char *name, int with_p, double beta, const double *beta_num, const T *p, T *pn,
                        const T *z, T *w, const T *r, T *rstore)
{
    const char *env = getenv("TEA_BEGIN_MARCH");
    if (env && env[0] == '0') return false;
    const Ext e = interior();
    const int nbands = (e.x1 - e.x0 + BEGIN_OUT_X) / BEGIN_OUT_X, ny = e.y1 - e.y0 + 1;
    const int target = cu_count() * 4 * stream_waves("TEA_BEGIN_WAVES", 8);
    int runs = target / nbands;
    runs = runs < 1 ? 1 : runs;
    int rows = (ny + runs - 1) / runs;
    rows = rows < 8 ? 8 : rows;
    runs = (ny + rows - 1) / rows;
    const int grid = (nbands * runs + MARCH_WAVES - 1) / MARCH_WAVES;
    if (grid > g.max_blocks) return false;
    g.last_grid = grid;
    klaunch(name, k_begin_march<T, STORE_R, Z_FROM_R>, dim3(grid), dim3(MARCH_WAVES * WAVE), e, pitch_of<T>(), rows, with_p,
            beta, beta_num, p, pn, z, w, (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), r, rstore, red_ctx());
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) return die("launch of %s failed: %s", name, hipGetErrorString(err));
    return true;
}

// z_src: where z is read (null: the z field); w_later: w not stored;
// store_r false: rstore = r and r.z left out (k_ppcg_end_rr formed them).
template <typename T>
bool ppcg_begin_launch(int with_p, double beta, const double *beta_num = nullptr, const T *z_src = nullptr,
                       bool w_later = false, bool store_r = true)
{
    const T *zf = z_src ? z_src : (const T *)FT<T>(X_Z);
    const long pt = pitch_of<T>();
    const long ring = 2L * (g.nx + 2L * g.h) * g.h + 2L * g.ny * g.h;
    bool &dirty = g.p_halo_dirty[std::is_same<T, double>::value ? 0 : 1];
    if (dirty) {  // the spare's halo ring takes p's (calc_p leaves the halo as it is)
        Timed t("p_halo_ring");
        hipLaunchKernelGGL(k_copy_ring<T>, dim3((unsigned)((ring + 255) / 256)), dim3(256), 0, g.stream,
                           FT<T>(X_P2), (const T *)FT<T>(F_P), pt, g.h, g.nx, g.ny);
        dirty = false;
    }
    bool ok;
    if (!store_r && begin_march_launch<T, false>(nm<T>("ppcg_begin_pw"), with_p, beta, beta_num,
                                                 (const T *)FT<T>(F_P), FT<T>(X_P2), zf,
                                                 w_later ? nullptr : FT<T>(X_W), (const T *)nullptr, (T *)nullptr))
        ok = true;   // else the begins below, whose rstore = r and r.z repeat the end's
    else if constexpr (std::is_same<T, float>::value)
        ok = begin_march_launch<T, true>("ppcg_begin_f32", with_p, beta, beta_num, (const T *)FT<T>(F_P), FT<T>(X_P2),
                                         zf, w_later ? nullptr : FT<T>(X_W), (const T *)FT<T>(F_R), FT<T>(X_RSTORE)) ||
             launch_x2("ppcg_begin_f32", k_ppcg_begin_x2, interior(), pt, with_p, beta, beta_num, (const T *)FT<T>(F_P),
                       FT<T>(X_P2), zf, w_later ? nullptr : FT<T>(X_W), (const T *)FT<T>(X_KX),
                       (const T *)FT<T>(X_KY), (const T *)FT<T>(F_R), FT<T>(X_RSTORE), red_ctx());
    else
        ok = begin_march_launch<T, true>(nm<T>("ppcg_begin"), with_p, beta, beta_num, (const T *)FT<T>(F_P), FT<T>(X_P2),
                                         zf, w_later ? nullptr : FT<T>(X_W), (const T *)FT<T>(F_R), FT<T>(X_RSTORE)) ||
             launch(nm<T>("ppcg_begin"), k_ppcg_begin<T>, interior(), pt, with_p, beta, beta_num, (const T *)FT<T>(F_P),
                    FT<T>(X_P2), zf, w_later ? nullptr : FT<T>(X_W), (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY),
                    (const T *)FT<T>(F_R), FT<T>(X_RSTORE), red_ctx());
    if (!ok) return false;
    if (std::is_same<T, double>::value) {
        std::swap(g.f[F_P], g.f[X_P2]);
        std::swap(g.fs[F_P], g.fs[X_P2]);
    } else {
        std::swap(g.fs32[F_P], g.fs32[X_P2]);
    }
    return true;
}

template <typename T> void ppcg_begin_impl(int with_p, double beta, double *out)
{
    if (ppcg_begin_launch<T>(with_p, beta)) fetch(out, 2);
}

// pw_rro: null (alpha given) or device memory holding p.w, r.z.
// defer: leave rtemp = r and utemp = sd to the next streaming inner pass
// (recorded in g.pending_rt_ut).
// w_from_p: w formed from p (block Jacobi; begin left it unwritten).
// from_r (with defer and w_from_p): z = M^-1 r and sd = z / theta are not
// stored either (g.pending_zr, g.pending_sd_r): the first inner pass forms sd
// from r, and z is written out only for a reader before the end replaces it.
template <typename T>
bool ppcg_ur_launch(double alpha, const double *pw_rro, double theta, bool defer = false, bool w_from_p = false,
                    bool from_r = false)
{
    T *rt = defer ? nullptr : FT<T>(X_RTEMP), *ut = defer ? nullptr : FT<T>(X_UTEMP);
    from_r = from_r && defer && w_from_p && g.prec == PREC_BLOCK;
    T *zf = from_r ? nullptr : FT<T>(X_Z), *sdf = from_r ? nullptr : FT<T>(F_SD);
    // z and sd as this ur leaves them (owed or stored): any older debt is void
    const int bits = 8 * (int)sizeof(T);
    g.pending_zr = g.pending_sd_r = 0;
    auto owe = [&](bool ok) {
        if (ok && from_r) {
            g.pending_zr = g.pending_sd_r = bits;
            g.sd_r_theta = theta;
        }
        return ok;
    };
    bool ok;
    if constexpr (std::is_same<T, float>::value) {
        if (w_from_p && g.prec == PREC_BLOCK) {
            ok = launch_x2_capped("k_ppcg_ur_sd_f32", 4096, k_ppcg_ur_sd_x2_wp, interior(), pitch_of<T>(), alpha,
                                  pw_rro, theta, FT<T>(F_U), (const T *)FT<T>(F_P), FT<T>(F_R), zf,
                                  (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), sdf, rt, ut, red_ctx());
            if (ok && defer) g.pending_rt_ut = 32;
            return owe(ok);
        }
        auto go = [&](auto kernel) {
            return launch_x2("k_ppcg_ur_sd_f32", kernel, interior(), pitch_of<T>(), alpha, pw_rro, theta, FT<T>(F_U),
                             (const T *)FT<T>(F_P), FT<T>(F_R), (const T *)FT<T>(X_W), FT<T>(X_Z),
                             (const T *)FT<T>(X_MI), (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), FT<T>(F_SD),
                             rt, ut, red_ctx());
        };
        ok = g.prec == PREC_BLOCK ? go(k_ppcg_ur_sd_x2<PREC_BLOCK>)
             : g.prec == PREC_DIAG ? go(k_ppcg_ur_sd_x2<PREC_DIAG>)
                                   : go(k_ppcg_ur_sd_x2<PREC_NONE>);
    } else if (w_from_p && g.prec == PREC_BLOCK) {
        ok = owe(launch("k_ppcg_ur_sd", k_ppcg_ur_sd_wp<T>, interior(), pitch_of<T>(), alpha, pw_rro, theta,
                        FT<T>(F_U), (const T *)FT<T>(F_P), FT<T>(F_R), zf, (const T *)FT<T>(X_KX),
                        (const T *)FT<T>(X_KY), sdf, rt, ut, red_ctx()));
    } else {
        ok = PREC_SWITCH_T(T, g.prec, "k_ppcg_ur_sd", k_ppcg_ur_sd, interior(), pitch_of<T>(), alpha, pw_rro, theta,
                           FT<T>(F_U), (const T *)FT<T>(F_P), FT<T>(F_R), (const T *)FT<T>(X_W), FT<T>(X_Z),
                           (const T *)FT<T>(X_MI), (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), FT<T>(F_SD),
                           rt, ut, red_ctx());
    }
    if (ok && defer) g.pending_rt_ut = 8 * (int)sizeof(T);
    return ok;
}

template <typename T> void ppcg_ur_impl(double alpha, double theta, double *rrn)
{
    if (ppcg_ur_launch<T>(alpha, nullptr, theta)) fetch(rrn, 1);
}

// begin, then ur with alpha = rro / pw formed on the device: one wait.
// z owed by a separate end (= utemp): begin reads utemp, ur rewrites z.
template <typename T> void ppcg_begin_ur_impl(int with_p, double beta, double theta, double *pw_rro, double *rrn)
{
    const bool wp = w_from_p<T>();
    settle_z_r();   // begin reads z
    const T *z_src = g.pending_z == 8 * (int)sizeof(T) ? (const T *)FT<T>(X_UTEMP) : nullptr;
    if (!ppcg_begin_launch<T>(with_p, beta, nullptr, z_src, wp) || !finish(0, 2, g.d_scal)) return;
    if (!ppcg_ur_launch<T>(0.0, g.d_scal, theta, use_march<T, 1>(), wp, sd_from_r<T>()) || !finish(2, 1, nullptr))
        return;
    g.pending_w = wp ? 8 * (int)sizeof(T) : 0;
    g.pending_z = 0;
    if (collect(pw_rro, 0, 2)) *rrn = g.h_out[2];
}

// end (update_z + calc_rrn), then begin_ur with beta = rrn_end / rro_prev
// formed on the device: one wait instead of two.  With a preconditioner, ur
// rewrites z, so z = utemp is not stored: begin reads utemp in its place.
// inner_steps > 0: inner steps 1..inner_steps are queued behind ur (passes as
// tea_engine_ppcg_inner_plan splits them) before the wait, which then waits
// for ur's values only; with a preconditioner the last pass also forms the
// next end's r.z (into slot 4 or 5), which the next call takes in place of
// its end when no other entry point came in between.
template <typename T>
void ppcg_end_begin_ur_impl(double rro_prev, double theta, double *rrn_end, double *pw_rro, double *rrn,
                            int inner_steps = 0)
{
    const bool z_from_utemp = g.prec != PREC_NONE;
    const T *ut = FT<T>(X_UTEMP);
    int eslot = 3;
    // end_rr: the end also stores r and forms the begin's r.z (k_ppcg_end_rr),
    // unless TEA_END_STORE_R=0: finished into h_out[6..7] and d_scal[0..1], where
    // begin takes beta's numerator and ur the r.z; begin's p.w then goes to slot 0
    const char *rr_env = getenv("TEA_END_STORE_R");
    const bool end_rr = !(g.end_slot >= 0 && g.end_call == g.calls - 1) && !(rr_env && rr_env[0] == '0');
    if (!end_rr && g.end_slot >= 0 && g.end_call == g.calls - 1) {
        eslot = g.end_slot;
    } else if (end_rr) {
        eslot = 6;
        if (!launch_capped(nm<T>("ppcg_end_rr"), 2048, k_ppcg_end_rr<T>, interior(), pitch_of<T>(), ut,
                           z_from_utemp ? nullptr : FT<T>(X_Z), (const T *)FT<T>(F_R), FT<T>(X_RSTORE), red_ctx()) ||
            !finish(6, 2, g.d_scal))
            return;
    } else if (!launch_capped(nm<T>("ppcg_end"), 2048, k_ppcg_end<T>, interior(), pitch_of<T>(), ut,
                              z_from_utemp ? nullptr : FT<T>(X_Z), (const T *)FT<T>(F_R),
                              (const T *)FT<T>(X_RSTORE), red_ctx()) ||
               !finish(3, 1, g.d_scal + 3)) {
        return;
    }
    g.end_slot = -1;
    g.pending_zr = 0;   // the end's z = utemp replaces z (begin reads utemp)
    const bool wp = w_from_p<T>();
    if (!ppcg_begin_launch<T>(1, rro_prev, end_rr ? g.d_scal : g.d_scal + eslot, z_from_utemp ? ut : nullptr, wp,
                              !end_rr) ||
        !finish(0, end_rr ? 1 : 2, g.d_scal))
        return;
    if (!ppcg_ur_launch<T>(0.0, g.d_scal, theta, use_march<T, 1>(), wp, sd_from_r<T>()) || !finish(2, 1, nullptr))
        return;
    g.pending_w = wp ? 8 * (int)sizeof(T) : 0;
    g.pending_z = 0;   // end recomputed what z owed; ur rewrote z
    if (inner_steps > 0) {
        if (hipEventRecord(g.ev_scal, g.stream) != hipSuccess) { die("hipEventRecord failed"); return; }
        for (int step = 1; step <= inner_steps;) {
            const int n = plan_first_pass(inner_steps - step + 1, pass_costs());
            const bool last = step + n - 1 == inner_steps;
            if (ppcg_inner_n_impl<T>(step - 1, n, last && z_from_utemp && fused_end<T>())) {
                const int ns = eslot == 4 ? 5 : 4;
                if (!finish(ns, 1, g.d_scal + ns)) return;
                g.end_slot = ns;
                g.end_call = g.calls;
            }
            step += n;
        }
        const hipError_t err = hipEventSynchronize(g.ev_scal);
        if (err != hipSuccess) { die("hipEventSynchronize failed: %s", hipGetErrorString(err)); return; }
        for (int j = 0; j < 2; j++) pw_rro[j] = g.h_out[j];
    } else if (!collect(pw_rro, 0, 2)) {
        return;
    }
    if (end_rr) pw_rro[1] = g.h_out[7];
    *rrn = g.h_out[2];
    *rrn_end = g.h_out[eslot];
}

// With a preconditioner z = utemp is owed rather than stored (the fused
// begin that follows reads utemp, and ur rewrites z).
template <typename T> void ppcg_end_impl(double *rrn)
{
    const bool owe_z = g.prec != PREC_NONE;
    if (!launch_capped(nm<T>("ppcg_end"), 2048, k_ppcg_end<T>, interior(), pitch_of<T>(), (const T *)FT<T>(X_UTEMP),
                       owe_z ? nullptr : FT<T>(X_Z), (const T *)FT<T>(F_R), (const T *)FT<T>(X_RSTORE), red_ctx()))
        return;
    g.pending_z = owe_z ? 8 * (int)sizeof(T) : 0;
    g.pending_zr = 0;   // z = utemp replaces z = M^-1 r
    fetch(rrn, 1);
}

// begin_ur, the inner steps in passes and end queued back to back with one
// host wait at the end: for callers that need only the end's r.z to go on.
template <typename T>
void ppcg_iteration_impl(int with_p, double beta, double theta, int inner_steps, double *pw_rro, double *rrn_ur,
                         double *rrn_end)
{
    const bool wp = w_from_p<T>();
    settle_z_r();   // begin reads z
    const T *z_src = g.pending_z == 8 * (int)sizeof(T) ? (const T *)FT<T>(X_UTEMP) : nullptr;
    if (!ppcg_begin_launch<T>(with_p, beta, nullptr, z_src, wp) || !finish(0, 2, g.d_scal)) return;
    if (!ppcg_ur_launch<T>(0.0, g.d_scal, theta, use_march<T, 1>(), wp, sd_from_r<T>()) || !finish(2, 1, nullptr))
        return;
    g.pending_w = wp ? 8 * (int)sizeof(T) : 0;
    g.pending_z = 0;
    const bool owe_z = g.prec != PREC_NONE;
    bool end_done = false;   // the last pass formed end's r.z (z = utemp then owed, as end leaves it)
    for (int step = 1; step <= inner_steps;) {
        const int n = plan_first_pass(inner_steps - step + 1, pass_costs());
        end_done = ppcg_inner_n_impl<T>(step - 1, n, step + n - 1 == inner_steps && owe_z && fused_end<T>());
        step += n;
    }
    settle_rt_ut();
    if (!end_done && !launch_capped(nm<T>("ppcg_end"), 2048, k_ppcg_end<T>, interior(), pitch_of<T>(),
                                    (const T *)FT<T>(X_UTEMP), owe_z ? nullptr : FT<T>(X_Z), (const T *)FT<T>(F_R),
                                    (const T *)FT<T>(X_RSTORE), red_ctx()))
        return;
    if (!finish(3, 1, nullptr)) return;
    g.pending_z = owe_z ? 8 * (int)sizeof(T) : 0;
    g.pending_zr = 0;   // z = utemp replaces z = M^-1 r
    if (!collect(pw_rro, 0, 2)) return;
    *rrn_ur = g.h_out[2];
    *rrn_end = g.h_out[3];
}

}  // namespace

#define BY_PRECISION(fn, ...) (g.fp32 ? fn<float>(__VA_ARGS__) : fn<double>(__VA_ARGS__))
#define REQUIRE_FP64()                                                                            \
    do { if (g.fp32) { die("%s: not available in the FP32 solver mode", __func__); return; } } while (0)

EXPORT void tea_leaf_cg_init_kernel_cuda_(double *rro)
{
    CHECK_INIT();
    REQUIRE_SOLVER(S_CG_FAMILY);
    BY_PRECISION(cg_init_impl, rro);
}

EXPORT void tea_leaf_cg_calc_w_kernel_cuda_(double *pw)
{
    CHECK_INIT();
    REQUIRE_SOLVER(S_CG_FAMILY);
    BY_PRECISION(cg_calc_w_impl, pw);
}

EXPORT void tea_leaf_cg_calc_ur_kernel_cuda_(double *alpha, double *rrn)
{
    CHECK_INIT();
    REQUIRE_SOLVER(S_CG_FAMILY);
    BY_PRECISION(cg_calc_ur_impl, *alpha, rrn);
}

EXPORT void tea_leaf_cg_calc_p_kernel_cuda_(double *beta)
{
    CHECK_INIT();
    REQUIRE_SOLVER(S_CG_FAMILY);
    BY_PRECISION(cg_calc_p_impl, *beta);
}

// ---- fused CG iteration (no reference counterpart) ---------------------------------
//
// tea_engine_cg_begin_ur: [cg_calc_p(beta) if with_p], cg_calc_w (pw),
// cg_calc_ur(alpha = rro / pw) (rrn), with identical results: calc_p is
// folded into calc_w (p' formed on the fly into the spare p buffer, w = A p'
// not stored), alpha is divided on the device, and ur forms w = A p' again
// from p' (one host wait, two passes instead of three).  w is then owed
// (pending_w): the next fused call rewrites it, any other entry point writes
// it out first.  FP32 mode runs the separate calls.  With jac_diag z = Mi r
// is owed too (pending_zr): ur does not store it, and the next fused begin
// forms it from r (zr) unless TEA_CG_Z_FROM_R=0.
namespace {
bool cg_z_from_r()
{
    const char *env = getenv("TEA_CG_Z_FROM_R"), *march = getenv("TEA_BEGIN_MARCH");
    return g.prec == PREC_DIAG && !(env && env[0] == '0') && !(march && march[0] == '0');
}

bool cg_begin_ur_f64(int with_p, double beta, double rro, double *pw, double *rrn, bool zr)
{
    const long pt = g.pitch;
    const long ring = 2L * (g.nx + 2L * g.h) * g.h + 2L * g.ny * g.h;
    if (g.p_halo_dirty[0]) {   // the spare's halo ring takes p's (calc_p leaves the halo as it is)
        Timed t("p_halo_ring");
        hipLaunchKernelGGL(k_copy_ring<double>, dim3((unsigned)((ring + 255) / 256)), dim3(256), 0, g.stream,
                           F(X_P2), (const double *)F(F_P), pt, g.h, g.nx, g.ny);
        g.p_halo_dirty[0] = false;
    }
    const double *src = g.prec != PREC_NONE ? (const double *)F(X_Z) : (const double *)F(F_R);
    bool begun = zr ? begin_march_launch<double, false, true>("cg_begin", with_p, beta, nullptr, (const double *)F(F_P),
                                                              F(X_P2), (const double *)F(F_R), nullptr, nullptr,
                                                              nullptr)
                    : begin_march_launch<double, false>("cg_begin", with_p, beta, nullptr, (const double *)F(F_P),
                                                        F(X_P2), src, nullptr, nullptr, nullptr);
    if (!begun) {
        if (zr) {   // the tile form reads z
            g.pending_zr = 64;
            settle_z_r();
        }
        if (!launch("cg_begin", k_ppcg_begin<double>, interior(), pt, with_p, beta, (const double *)nullptr,
                    (const double *)F(F_P), F(X_P2), src, (double *)nullptr, (const double *)F(X_KX),
                    (const double *)F(X_KY), (const double *)nullptr, (double *)nullptr, red_ctx()))
            return false;
    }
    std::swap(g.f[F_P], g.f[X_P2]);
    std::swap(g.fs[F_P], g.fs[X_P2]);
    if (!finish(0, 1, g.d_scal)) return false;
    const bool owe_z = cg_z_from_r();
    auto go = [&](auto kernel) {
        return launch("cg_ur", kernel, interior(), pt, rro, (const double *)g.d_scal, F(F_U), (const double *)F(F_P),
                      F(F_R), owe_z ? (double *)nullptr : F(X_Z), (const double *)F(X_MI), (const double *)F(X_KX),
                      (const double *)F(X_KY), red_ctx());
    };
    const bool ok = g.prec == PREC_BLOCK ? go(k_cg_ur_wp<double, PREC_BLOCK>)
                    : g.prec == PREC_DIAG ? go(k_cg_ur_wp<double, PREC_DIAG>)
                                          : go(k_cg_ur_wp<double, PREC_NONE>);
    if (!ok) return false;
    g.pending_zr = owe_z ? 64 : 0;
    if (!finish(2, 1, nullptr)) return false;
    g.pending_w = 64;
    if (!collect(pw, 0, 1)) return false;
    *rrn = g.h_out[2];
    return true;
}
}  // namespace

EXPORT void tea_engine_cg_begin_ur(int with_p, double beta, double rro, double *pw, double *rrn)
{
    CHECK_INIT_KEEP_PENDING();
    REQUIRE_SOLVER(S_CG_FAMILY);
    if (g.fp32) {   // the composition, call by call
        settle_pending();
        if (with_p) cg_calc_p_impl<float>(beta);
        cg_calc_w_impl<float>(pw);
        cg_calc_ur_impl<float>(rro / *pw, rrn);
        return;
    }
    g.pending_w = 0;   // rewritten below
    // z = Mi r owed by the previous fused iteration: this begin forms it from r
    const bool zr = g.pending_zr == 64 && g.prec == PREC_DIAG && cg_z_from_r();
    if (zr) g.pending_zr = 0;
    settle_pending();
    cg_begin_ur_f64(with_p, beta, rro, pw, rrn, zr);
}

namespace {
void upload_coefs(const double *a, const double *b, int n)
{
    if (n < 0) n = 0;
    g.ch_a.assign(a, a + n);
    g.ch_b.assign(b, b + n);
}

template <typename T, int FIRST>
bool cheby_p(double theta, double alpha, double beta)
{
    const char *name = nm<T>(FIRST ? "cheby_init_p" : "cheby_calc_p");
    const Ext in = interior();
    auto args = [&](auto kernel) {
        return launch(name, kernel, in, pitch_of<T>(), (const T *)FT<T>(F_U), (const T *)FT<T>(X_U0), FT<T>(F_P),
                      FT<T>(F_R), FT<T>(X_W), (const T *)FT<T>(X_MI), (const T *)FT<T>(X_KX),
                      (const T *)FT<T>(X_KY), theta, alpha, beta);
    };
    if (g.prec == PREC_BLOCK) return args(k_cheby_p<T, PREC_BLOCK, FIRST>);
    if (g.prec == PREC_DIAG) return args(k_cheby_p<T, PREC_DIAG, FIRST>);
    return args(k_cheby_p<T, PREC_NONE, FIRST>);
}

template <typename T> void cheby_init_impl(double theta)
{
    if (!cheby_p<T, 1>(theta, 0.0, 0.0)) return;
    launch(nm<T>("cheby_calc_u"), k_add<T>, interior(), pitch_of<T>(), FT<T>(F_U), (const T *)FT<T>(F_P));
}

// One Chebyshev iteration.  Without block Jacobi (and unless
// TEA_FUSED_CHEBY=0) in one streaming pass (k_cheby_march): u + p goes to the
// spare u buffer, which becomes u, and w, r of the previous u (kept in the
// spare) are owed (pending_cheby): the next iteration recomputes them, any
// other entry point writes them out first.
template <typename T> void cheby_iterate_impl(double alpha, double beta)
{
    const char *env = getenv("TEA_FUSED_CHEBY");
    if (g.prec != PREC_BLOCK && !(env && env[0] == '0')) {
        const Ext e = interior();
        const long pt = pitch_of<T>();
        const long ring = 2L * (g.nx + 2L * g.h) * g.h + 2L * g.ny * g.h;
        if (g.u_ring_dirty) {   // the spare takes u's halo, which u + p keeps
            Timed t("u_halo_ring");
            hipLaunchKernelGGL(k_copy_ring<T>, dim3((unsigned)((ring + 255) / 256)), dim3(256), 0, g.stream,
                               FT<T>(X_U2), (const T *)FT<T>(F_U), pt, g.h, g.nx, g.ny);
            g.u_ring_dirty = false;
        }
        const int nbands = (e.x1 - e.x0 + BEGIN_OUT_X) / BEGIN_OUT_X, ny = e.y1 - e.y0 + 1;
        int runs = cu_count() * 4 * stream_waves("TEA_CHEBY_WAVES", 4) / nbands;
        runs = runs < 1 ? 1 : runs;
        int rows = (ny + runs - 1) / runs;
        rows = rows < 8 ? 8 : rows;
        runs = (ny + rows - 1) / rows;
        const unsigned grid = (unsigned)((nbands * runs + MARCH_WAVES - 1) / MARCH_WAVES);
        auto go = [&](auto kernel) {
            klaunch(nm<T>("cheby_iterate"), kernel, dim3(grid), dim3(MARCH_WAVES * WAVE), e, pt, rows,
                    (const T *)FT<T>(F_U), FT<T>(X_U2), (const T *)FT<T>(X_U0), FT<T>(F_P), (const T *)FT<T>(X_MI),
                    (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), alpha, beta);
        };
        if (g.prec == PREC_DIAG) go(k_cheby_march<T, PREC_DIAG>);
        else go(k_cheby_march<T, PREC_NONE>);
        const hipError_t err = hipGetLastError();
        if (err != hipSuccess) { die("launch of k_cheby_march failed: %s", hipGetErrorString(err)); return; }
        if (std::is_same<T, double>::value) {
            std::swap(g.f[F_U], g.f[X_U2]);
            std::swap(g.fs[F_U], g.fs[X_U2]);
        } else {
            std::swap(g.fs32[F_U], g.fs32[X_U2]);
        }
        g.pending_cheby = 8 * (int)sizeof(T);
        return;
    }
    if (!cheby_p<T, 0>(0.0, alpha, beta)) return;
    launch(nm<T>("cheby_calc_u"), k_add<T>, interior(), pitch_of<T>(), FT<T>(F_U), (const T *)FT<T>(F_P));
}

// Chebyshev iterations per streaming pass (k_cheby_ring): the most whose
// pipeline fits 256 registers at two waves per SIMD in FP32 and one in FP64.
template <typename T> constexpr int cheby_pass_max() { return sizeof(T) == 4 ? 4 : 3; }

// Chebyshev iterations k0+1 .. k0+n (0-based coefficient index k0) as
// streaming passes of up to cheby_pass_max iterations (TEA_CHEBY_PASSES=0:
// one cheby_iterate_impl each); fields of 4 GiB or more take the latter.
template <typename T> void cheby_iterate_n_impl(int k0, int n)
{
    const char *env = getenv("TEA_CHEBY_PASSES");
    const bool fits = (double)pitch_of<T>() * (double)g.rows * sizeof(T) < 4294967296.0 - 65536.0;
    if ((env && env[0] == '0') || !fits) {
        for (int k = k0; k < k0 + n; k++) cheby_iterate_impl<T>(g.ch_a[k], g.ch_b[k]);
        return;
    }
    const Ext e = interior();
    const long pt = pitch_of<T>();
    const long ring = 2L * (g.nx + 2L * g.h) * g.h + 2L * g.ny * g.h;
    for (int k = k0; k < k0 + n;) {
        const int ns = std::min(cheby_pass_max<T>(), k0 + n - k);
        if (g.u_ring_dirty) {   // the spare takes u's halo, which the pass keeps
            Timed t("u_halo_ring");
            hipLaunchKernelGGL(k_copy_ring<T>, dim3((unsigned)((ring + 255) / 256)), dim3(256), 0, g.stream,
                               FT<T>(X_U2), (const T *)FT<T>(F_U), pt, g.h, g.nx, g.ny);
            g.u_ring_dirty = false;
        }
        bool &p_dirty = g.p_halo_dirty[std::is_same<T, double>::value ? 0 : 1];
        if (p_dirty) {   // and p's, which the pass keeps as well
            Timed t("p_halo_ring");
            hipLaunchKernelGGL(k_copy_ring<T>, dim3((unsigned)((ring + 255) / 256)), dim3(256), 0, g.stream,
                               FT<T>(X_P2), (const T *)FT<T>(F_P), pt, g.h, g.nx, g.ny);
            p_dirty = false;
        }
        StepCoefs cf{};
        for (int j = 0; j < ns; j++) {
            cf.a[j] = g.ch_a[k + j];
            cf.b[j] = g.ch_b[k + j];
        }
        auto go = [&](auto ns_c) {
            constexpr int NS = decltype(ns_c)::value;
            constexpr int OX = march_out_x(NS);
            const int nbands = (e.x1 - e.x0 + OX) / OX, nstrips = (e.y1 - e.y0 + STRIP) / STRIP;
            const int per_simd = sizeof(T) == 4 ? 2 : 1;
            int nchunks = cu_count() * 4 * per_simd / nbands;
            nchunks = nchunks < 1 ? 1 : (nchunks > nstrips ? nstrips : nchunks);
            int chunk = (nstrips + nchunks - 1) / nchunks;
            if (const char *c = getenv("TEA_MARCH_CHUNK")) {
                const int v = atoi(c);
                if (v > 0) chunk = v < nstrips ? v : nstrips;
            }
            nchunks = (nstrips + chunk - 1) / chunk;
            const unsigned grid = (unsigned)((nbands * nchunks + MARCH_WAVES - 1) / MARCH_WAVES);
            auto launch_prec = [&](auto kernel) {
                klaunch(nm<T>("cheby_pass"), kernel, dim3(grid), dim3(MARCH_WAVES * WAVE), e, pt, chunk,
                        (const T *)FT<T>(F_U), FT<T>(X_U2), (const T *)FT<T>(F_P), FT<T>(X_P2),
                        (const T *)FT<T>(X_U0), FT<T>(F_R), FT<T>(X_W), (const T *)FT<T>(X_MI),
                        (const T *)FT<T>(X_KX), (const T *)FT<T>(X_KY), cf);
            };
            if (g.prec == PREC_BLOCK) launch_prec(k_cheby_ring<T, PREC_BLOCK, NS>);
            else if (g.prec == PREC_DIAG) launch_prec(k_cheby_ring<T, PREC_DIAG, NS>);
            else launch_prec(k_cheby_ring<T, PREC_NONE, NS>);
        };
        switch (ns) {
        case 1: go(std::integral_constant<int, 1>{}); break;
        case 2: go(std::integral_constant<int, 2>{}); break;
        case 3: go(std::integral_constant<int, 3>{}); break;
        default: go(std::integral_constant<int, cheby_pass_max<T>() >= 4 ? 4 : 3>{}); break;
        }
        const hipError_t err = hipGetLastError();
        if (err != hipSuccess) { die("launch of k_cheby_ring failed: %s", hipGetErrorString(err)); return; }
        if (std::is_same<T, double>::value) {
            std::swap(g.f[F_U], g.f[X_U2]);
            std::swap(g.fs[F_U], g.fs[X_U2]);
            std::swap(g.f[F_P], g.f[X_P2]);
            std::swap(g.fs[F_P], g.fs[X_P2]);
        } else {
            std::swap(g.fs32[F_U], g.fs32[X_U2]);
            std::swap(g.fs32[F_P], g.fs32[X_P2]);
        }
        g.pending_cheby = 0;   // the pass wrote the last iteration's w and r
        k += ns;               // both u buffers keep u's halo ring: the spare stays clean
    }
}
}  // namespace

EXPORT void tea_leaf_cheby_init_kernel_cuda_(const double *ch_alphas, const double *ch_betas, int *n_coefs,
                                             const double *theta)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_CHEBY);
    upload_coefs(ch_alphas, ch_betas, *n_coefs);
    BY_PRECISION(cheby_init_impl, *theta);
}

EXPORT void tea_leaf_cheby_iterate_kernel_cuda_(const int *cheby_calc_step)
{
    CHECK_INIT_KEEP_PENDING();
    g.pending_cheby = 0;   // this iteration forms w and r again
    settle_pending();
    REQUIRE_SOLVER(1 << SOLVER_CHEBY);
    const int k = *cheby_calc_step - 1;
    if (k < 0 || k >= (int)g.ch_a.size()) {
        die("cheby_iterate: step %d outside the %d uploaded coefficients", k + 1, (int)g.ch_a.size());
        return;
    }
    BY_PRECISION(cheby_iterate_impl, g.ch_a[k], g.ch_b[k]);
}

// Chebyshev iterations step .. step + nsteps - 1 (1-based coefficient index,
// as tea_leaf_cheby_iterate_kernel_cuda_ takes it) in streaming passes.
EXPORT void tea_engine_cheby_iterate_n(int step, int nsteps)
{
    CHECK_INIT_KEEP_PENDING();
    g.pending_cheby = 0;   // the passes form w and r again
    settle_pending();
    REQUIRE_SOLVER(1 << SOLVER_CHEBY);
    const int k = step - 1;
    if (nsteps < 1 || k < 0 || k + nsteps > (int)g.ch_a.size()) {
        die("cheby_iterate_n: steps %d..%d outside the %d uploaded coefficients", step, step + nsteps - 1,
            (int)g.ch_a.size());
        return;
    }
    BY_PRECISION(cheby_iterate_n_impl, k, nsteps);
}

EXPORT void tea_leaf_jacobi_solve_kernel_cuda_(double *error)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_JACOBI);
    REQUIRE_FP64();
    const Ext in = interior();
    launch("jacobi_copy_u", k_copy<double>, in, g.pitch, F(X_MI), (const double *)F(F_U));
    if (!launch("jacobi_solve", k_jacobi, in, g.pitch, (const double *)F(X_KX), (const double *)F(X_KY),
                (const double *)F(X_U0), F(F_U), (const double *)F(X_MI), red_ctx()))
        return;
    fetch(error, 1);
}

EXPORT void tea_leaf_ppcg_init_constants_cuda_(const double *ch_alphas, const double *ch_betas, int *n_inner_steps)
{
    CHECK_INIT();
    upload_coefs(ch_alphas, ch_betas, *n_inner_steps);
}

EXPORT void tea_leaf_ppcg_init_kernel_cuda_(const int *step, double *rro)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    const int st = *step;
    if (st < 1 || st > 3) { die("ppcg_init: invalid step %d", st); return; }
    BY_PRECISION(ppcg_init_impl, st, rro);
}

// Both overwrite sd, rtemp and utemp: rtemp and utemp still owed need no writing.
static void drop_or_settle_pending()
{
    g.pending_sd_r = 0;   // sd is rewritten
    settle_z();   // init_sd_new reads z
    if (g.pending_rt_ut == (g.fp32 ? 32 : 64)) g.pending_rt_ut = 0;
    else settle_rt_ut();
}

// The fused begin rewrites w: an owed w is dropped, owed rtemp and utemp are
// written out.
static void drop_w_settle_rest()
{
    g.pending_w = 0;
    settle_rt_ut();
}

EXPORT void tea_leaf_ppcg_init_sd_kernel_cuda_(const double *theta)
{
    CHECK_INIT_KEEP_PENDING();
    drop_or_settle_pending();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    BY_PRECISION(ppcg_init_sd_impl, *theta);
}

EXPORT void tea_leaf_ppcg_init_sd_new_kernel_cuda_(const double *theta)
{
    CHECK_INIT_KEEP_PENDING();
    drop_or_settle_pending();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    BY_PRECISION(ppcg_init_sd_new_impl, *theta);
}

EXPORT void tea_leaf_ppcg_inner_kernel_cuda_(int *ppcg_cur_step, int *bounds_extra, int *chunk_neighbours)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    const int k = *ppcg_cur_step - 1;
    if (k < 0 || k >= (int)g.ch_a.size()) {
        die("ppcg_inner: step %d outside the %d uploaded coefficients", k + 1, (int)g.ch_a.size());
        return;
    }
    const int be = *bounds_extra;
    if (be < 0 || be >= g.h) { die("ppcg_inner: invalid bounds_extra %d", be); return; }
    // matrix-powers extent, tea_leaf_kernel_cuda.cu:418-452
    const int h = g.h;
    Ext e;
    e.x0 = (chunk_neighbours[TEA_CHUNK_LEFT - 1] == TEA_EXTERNAL_FACE) ? h : h - be;
    e.x1 = h + g.nx - 1 + ((chunk_neighbours[TEA_CHUNK_RIGHT - 1] == TEA_EXTERNAL_FACE) ? 0 : be);
    e.y0 = (chunk_neighbours[TEA_CHUNK_BOTTOM - 1] == TEA_EXTERNAL_FACE) ? h : h - be;
    e.y1 = h + g.ny - 1 + ((chunk_neighbours[TEA_CHUNK_TOP - 1] == TEA_EXTERNAL_FACE) ? 0 : be);
    BY_PRECISION(ppcg_inner_impl, e, g.ch_a[k], g.ch_b[k]);
}

EXPORT void tea_leaf_ppcg_store_r_kernel_cuda_(void)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    BY_PRECISION(ppcg_store_r_impl);
}

EXPORT void tea_leaf_ppcg_update_z_kernel_cuda_(void)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    BY_PRECISION(ppcg_update_z_impl);
}

EXPORT void tea_leaf_ppcg_calc_rrn_kernel_cuda_(double *norm)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    BY_PRECISION(ppcg_calc_rrn_impl, norm);
}

EXPORT void tea_leaf_ppcg_calc_2norm_kernel_cuda_(double *norm)
{
    CHECK_INIT();
    BY_PRECISION(ppcg_calc_2norm_impl, norm);
}

EXPORT void tea_leaf_ppcg_calc_p_kernel_cuda_(double *beta)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    BY_PRECISION(ppcg_calc_p_impl, *beta);
}

// =============================================================================
// Extensions
// =============================================================================

// Fused PPCG outer iteration (no reference counterpart): the same arithmetic as
// the entry points named in include/tea_leaf_engine.h, in fewer passes.
EXPORT void tea_engine_ppcg_begin(int with_p, double beta, double *pw_rro)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    BY_PRECISION(ppcg_begin_impl, with_p, beta, pw_rro);
}

EXPORT void tea_engine_ppcg_ur(double alpha, double theta, double *rrn)
{
    CHECK_INIT();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    BY_PRECISION(ppcg_ur_impl, alpha, theta, rrn);
}

EXPORT void tea_engine_ppcg_begin_ur(int with_p, double beta, double theta, double *pw_rro, double *rrn)
{
    CHECK_INIT_KEEP_PENDING();
    drop_w_settle_rest();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    BY_PRECISION(ppcg_begin_ur_impl, with_p, beta, theta, pw_rro, rrn);
}

EXPORT void tea_engine_ppcg_end_begin_ur(double rro_prev, double theta, double *rrn_end, double *pw_rro,
                                         double *rrn)
{
    CHECK_INIT_KEEP_PENDING();
    drop_w_settle_rest();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    BY_PRECISION(ppcg_end_begin_ur_impl, rro_prev, theta, rrn_end, pw_rro, rrn);
}

EXPORT void tea_engine_ppcg_end_begin_ur_inner(double rro_prev, double theta, int inner_steps, double *rrn_end,
                                               double *pw_rro, double *rrn)
{
    CHECK_INIT_KEEP_PENDING();
    drop_w_settle_rest();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    if (inner_steps < 0 || inner_steps > (int)g.ch_a.size()) {
        die("ppcg_end_begin_ur_inner: %d inner steps, %d coefficients uploaded", inner_steps, (int)g.ch_a.size());
        return;
    }
    BY_PRECISION(ppcg_end_begin_ur_impl, rro_prev, theta, rrn_end, pw_rro, rrn, inner_steps);
}

EXPORT void tea_engine_ppcg_end(double *rrn)
{
    CHECK_INIT_KEEP_PENDING();   // reads neither p nor w: an owed w stays owed
    settle_rt_ut();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    BY_PRECISION(ppcg_end_impl, rrn);
}

EXPORT void tea_engine_ppcg_iteration(int with_p, double beta, double theta, int inner_steps, double *pw_rro,
                                      double *rrn_ur, double *rrn_end)
{
    CHECK_INIT_KEEP_PENDING();
    drop_w_settle_rest();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    if (inner_steps < 0 || inner_steps > (int)g.ch_a.size()) {
        die("ppcg_iteration: %d inner steps, %d coefficients uploaded", inner_steps, (int)g.ch_a.size());
        return;
    }
    BY_PRECISION(ppcg_iteration_impl, with_p, beta, theta, inner_steps, pw_rro, rrn_ur, rrn_end);
}

// Inner steps `step` .. `step + nsteps - 1` (1-based, nsteps 1..5) over the
// interior in one pass: tea_leaf_ppcg_inner_kernel_cuda_ once per step with
// bounds_extra 0 and every face external (one chunk).
EXPORT void tea_engine_ppcg_inner_n(int step, int nsteps)
{
    CHECK_INIT_KEEP_PENDING();   // a streaming pass takes rtemp and utemp still owed from r and sd
    settle_z_utemp();            // the pass rewrites utemp (r, which z = M^-1 r reads, it keeps)
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    const int k = step - 1;
    if (nsteps < 1 || nsteps > MAX_PASS_STEPS) {
        die("ppcg_inner_n: %d steps in one pass (1..%d)", nsteps, MAX_PASS_STEPS);
        return;
    }
    if (k < 0 || k + nsteps > (int)g.ch_a.size()) {
        die("ppcg_inner_n: steps %d..%d outside the %d uploaded coefficients", step, step + nsteps - 1,
            (int)g.ch_a.size());
        return;
    }
    BY_PRECISION(ppcg_inner_n_impl, k, nsteps);
}

EXPORT void tea_engine_ppcg_inner2(int step) { tea_engine_ppcg_inner_n(step, 2); }

// Steps the next pass should take out of `remaining` consecutive inner steps
// (the first pass of the cheapest plan for the current precision mode).
EXPORT int tea_engine_ppcg_inner_plan(int remaining)
{
    return plan_first_pass(remaining, pass_costs());
}

EXPORT void tea_engine_set_deck_path(const char *path) { g_deck_path = path ? path : "tea.in"; }
EXPORT void tea_engine_set_error_mode(int mode) { g_error_mode = mode; }
EXPORT int tea_engine_last_error(char *buf, int buflen)
{
    if (buf && buflen > 0) snprintf(buf, (size_t)buflen, "%s", g_error);
    return g_error[0] != 0;
}
EXPORT void tea_engine_clear_error(void) { g_error[0] = 0; }
EXPORT void tea_engine_info(int *nx, int *ny, int *halo_depth, int *solver, int *preconditioner)
{
    *nx = g.nx;
    *ny = g.ny;
    *halo_depth = g.h;
    *solver = g.solver;
    *preconditioner = g.prec;
}

EXPORT int tea_engine_get_field(int id, double *host)
{
    if (g.initialised) settle_pending();   // rtemp and utemp owed by tea_engine_ppcg_begin_ur
    if (!g.initialised || id < 1 || id >= X_COUNT) return -1;
    const size_t w = (size_t)g.nx + 2 * g.h;
    HIP_OK_R(hipMemcpy2DAsync(host, w * sizeof(double), g.f[id] + (g.lead - g.h), g.pitch * sizeof(double),
                              w * sizeof(double), (size_t)g.rows, hipMemcpyDeviceToHost, g.stream), -1);
    HIP_OK_R(hipStreamSynchronize(g.stream), -1);
    return 0;
}

EXPORT int tea_engine_set_field(int id, const double *host)
{
    if (g.initialised) settle_pending();   // rtemp and utemp owed by tea_engine_ppcg_begin_ur
    if (!g.initialised || id < 1 || id >= X_COUNT) return -1;
    if (id == F_SD) g.sd_halo_dirty = true;
    if (id == F_P) g.p_halo_dirty[0] = true;
    if (id == X_RTEMP) g.rt_halo_dirty[0] = true;
    g.u_ring_dirty = true;
    const size_t w = (size_t)g.nx + 2 * g.h;
    HIP_OK_R(hipMemcpy2DAsync(g.f[id] + (g.lead - g.h), g.pitch * sizeof(double), host, w * sizeof(double),
                              w * sizeof(double), (size_t)g.rows, hipMemcpyHostToDevice, g.stream), -1);
    HIP_OK_R(hipStreamSynchronize(g.stream), -1);
    return 0;
}

EXPORT void tea_engine_profile_enable(int on)
{
    if (!on && g.stream) {
        (void)hipStreamSynchronize(g.stream);
        g.prof.drain(true);
    }
    if (on && !g.prof.on) g.prof.reserve(1 << 15);
    g.prof.on = on != 0;
}
// Times only the launches whose name starts with `prefix` (NULL or "": all);
// each timed launch costs the run a few microseconds.
EXPORT void tea_engine_profile_only(const char *prefix) { g.prof.only = prefix ? prefix : ""; }
// Times one in `every` (>= 1) of each name's wanted launches, from the first.
EXPORT void tea_engine_profile_sample(int every) { g.prof.every = every > 1 ? every : 1; }
EXPORT void tea_engine_profile_reset(void)
{
    if (g.stream) (void)hipStreamSynchronize(g.stream);
    g.prof.drain(true);
    for (auto &c : g.prof.counts) c = 0;
    for (auto &c : g.prof.seen) c = 0;
    for (auto &m : g.prof.ms) m = 0.0;
}
EXPORT int tea_engine_profile_get(const char *kernel, long *count, double *total_ms)
{
    if (g.stream) (void)hipStreamSynchronize(g.stream);
    g.prof.drain(true);
    auto it = g.prof.ids.find(kernel);
    if (it == g.prof.ids.end()) {
        *count = 0;
        *total_ms = 0.0;
        return -1;
    }
    *count = g.prof.counts[it->second];
    *total_ms = g.prof.ms[it->second];
    return 0;
}
EXPORT int tea_engine_profile_names(char *buf, int buflen)
{
    std::string all;
    for (auto &n : g.prof.names) all += n + "\n";
    if (buf && buflen > 0) snprintf(buf, (size_t)buflen, "%s", all.c_str());
    return (int)all.size();
}
EXPORT void *tea_engine_stream(void) { return (void *)g.stream; }
EXPORT void tea_engine_synchronize(void)
{
    if (g.stream) (void)hipStreamSynchronize(g.stream);
    if (g.prof.on) g.prof.drain(false);
}
EXPORT void tea_engine_finalize(void) { release_all(); }

// ---- mixed precision -------------------------------------------------------------

EXPORT int tea_engine_set_precision(int bits)
{
    if (g.initialised) settle_pending();   // rtemp and utemp owed by tea_engine_ppcg_begin_ur
    if (!g.initialised) { die("%s called before initialise_cuda_", __func__); return -1; }
    if (bits != 32 && bits != 64) { die("tea_engine_set_precision: %d bits is not 32 or 64", bits); return -1; }
    if (bits == 32 && !g.slab32) {
        g.lead32 = (g.h + 31) / 32 * 32;                       // 128-byte aligned interior
        g.pitch32 = (long)((g.lead32 + g.nx + g.h + 31) / 32 * 32);
        const size_t elems = ((size_t)g.pitch32 * (size_t)g.rows + 32 + 63) / 64 * 64;
        HIP_OK_R(hipMalloc(&g.slab32, elems * N_SLOTS * sizeof(float)), -1);
        HIP_OK_R(hipMemsetAsync(g.slab32, 0, elems * N_SLOTS * sizeof(float), g.stream), -1);
        for (int k = 1; k < N_SLOTS; k++) g.fs32[k] = g.slab32 + elems * k + (g.lead32 - g.h);
    }
    g.fp32 = bits == 32;
    return 0;
}

// Leave the FP32 mode at the end of an FP32 solve: tea_engine_set_precision(64),
// except that the FP32 inner-solver fields still owed (w, z, sd, rtemp,
// utemp) are dropped unwritten.  The next FP32 solve writes each of them
// before reading it; until then their FP32 copies hold unspecified values.
EXPORT int tea_engine_mp_leave(void)
{
    if (g.initialised) {
        for (int *debt : {&g.pending_w, &g.pending_z, &g.pending_zr, &g.pending_sd_r, &g.pending_rt_ut})
            if (*debt == 32) *debt = 0;
    }
    return tea_engine_set_precision(64);
}

EXPORT void tea_engine_mp_setup(void)
{
    CHECK_INIT();
    if (!g.slab32) { die("tea_engine_mp_setup: call tea_engine_set_precision(32) first"); return; }
    launch("mp_coef", k_mp_coef, Ext{0, g.nx + 2 * g.h - 1, 0, g.ny + 2 * g.h - 1}, g.pitch, g.pitch32,
           (const double *)F(X_KX), (const double *)F(X_KY), g.fs32[X_KX], g.fs32[X_KY]);
}

EXPORT void tea_engine_mp_begin(double *rr)
{
    CHECK_INIT();
    if (!g.slab32) { die("tea_engine_mp_begin: call tea_engine_set_precision(32) first"); return; }
    if (!launch("mp_begin", k_mp_begin, interior(), g.pitch, g.pitch32, (const double *)F(F_R), g.fs32[F_R],
                g.fs32[X_U0], g.fs32[F_U], red_ctx()))
        return;
    fetch(rr, 1);
}

EXPORT void tea_engine_mp_end(void)
{
    CHECK_INIT();
    if (!g.slab32) { die("tea_engine_mp_end: call tea_engine_set_precision(32) first"); return; }
    launch("mp_end", k_mp_end, interior(), g.pitch, g.pitch32, F(F_U), (const float *)g.fs32[F_U]);
}

// mp_end, set_precision(64), calc_residual, calc_2norm(1) -> rr64,
// set_precision(32), mp_begin -> rr32, ppcg_init step 2, init_sd_new(theta):
// one pass (k_mp_replace, u' into the spare u buffer, which becomes u), the
// u32 = 0 fill, one wait.  rtemp and utemp still owed need no writing.  With
// block Jacobi and the inner passes forming sd from r, z, sd, rtemp and utemp
// are left owed from r as a fused ur leaves them (pending_zr, pending_sd_r,
// pending_rt_ut; TEA_SD_FROM_R=0 writes them here).
EXPORT void tea_engine_mp_replace(double theta, double *rr64, double *rr32)
{
    CHECK_INIT_KEEP_PENDING();
    REQUIRE_SOLVER(1 << SOLVER_PPCG);
    if (!g.fp32 || !g.slab32) { die("tea_engine_mp_replace: call from the FP32 solver mode"); return; }
    g.pending_zr = 0;   // z (a block-Jacobi debt) is rewritten with r
    drop_or_settle_pending();
    const Ext in = interior();
    const long ring = 2L * (g.nx + 2L * g.h) * g.h + 2L * g.ny * g.h;
    {   // the spare takes u's halo, which u' keeps
        Timed t("u_halo_ring");
        hipLaunchKernelGGL(k_copy_ring<double>, dim3((unsigned)((ring + 255) / 256)), dim3(256), 0, g.stream, F(X_U2),
                           (const double *)F(F_U), g.pitch, g.h, g.nx, g.ny);
    }
    const bool owe = g.prec == PREC_BLOCK && sd_from_r<float>();
    auto go = [&](auto kernel) {
        return launch(owe ? "mp_replace_r" : "mp_replace", kernel, in, g.pitch, g.pitch32, (const double *)F(F_U),
                      F(X_U2), (const double *)F(X_U0), F(F_R), (const double *)F(X_KX), (const double *)F(X_KY),
                      (const float *)g.fs32[F_U], g.fs32[F_R], g.fs32[X_U0], g.fs32[X_Z], g.fs32[F_SD],
                      g.fs32[X_RTEMP], g.fs32[X_UTEMP], (const float *)g.fs32[X_MI], (const float *)g.fs32[X_KX],
                      (const float *)g.fs32[X_KY], theta, red_ctx());
    };
    const bool ok = owe ? go(k_mp_replace<PREC_BLOCK, true>)
                    : g.prec == PREC_BLOCK ? go(k_mp_replace<PREC_BLOCK>)
                    : g.prec == PREC_DIAG  ? go(k_mp_replace<PREC_DIAG>)
                                           : go(k_mp_replace<PREC_NONE>);
    if (!ok || !finish(0, 2, nullptr)) return;
    if (owe) {
        g.pending_zr = g.pending_sd_r = g.pending_rt_ut = 32;
        g.sd_r_theta = theta;
    }
    std::swap(g.f[F_U], g.f[X_U2]);
    std::swap(g.fs[F_U], g.fs[X_U2]);
    if (!launch("mp_zero_u32", k_fill<float>, in, g.pitch32, g.fs32[F_U], 0.0f)) return;
    double v[2];
    if (!collect(v, 0, 2)) return;
    *rr64 = v[0];
    *rr32 = v[1];
}

EXPORT int tea_engine_get_field32(int id, float *host)
{
    if (g.initialised) settle_pending();   // rtemp and utemp owed by tea_engine_ppcg_begin_ur
    if (!g.initialised || !g.slab32 || id < 1 || id >= X_COUNT) return -1;
    const size_t w = (size_t)g.nx + 2 * g.h;
    HIP_OK_R(hipMemcpy2DAsync(host, w * sizeof(float), g.fs32[id], g.pitch32 * sizeof(float), w * sizeof(float),
                              (size_t)g.rows, hipMemcpyDeviceToHost, g.stream), -1);
    HIP_OK_R(hipStreamSynchronize(g.stream), -1);
    return 0;
}

EXPORT int tea_engine_set_field32(int id, const float *host)
{
    if (g.initialised) settle_pending();   // rtemp and utemp owed by tea_engine_ppcg_begin_ur
    g.u_ring_dirty = true;
    if (!g.initialised || !g.slab32 || id < 1 || id >= X_COUNT) return -1;
    const size_t w = (size_t)g.nx + 2 * g.h;
    HIP_OK_R(hipMemcpy2DAsync(g.fs32[id], g.pitch32 * sizeof(float), host, w * sizeof(float), w * sizeof(float),
                              (size_t)g.rows, hipMemcpyHostToDevice, g.stream), -1);
    if (id == F_SD) g.sd32_halo_dirty = true;
    if (id == F_P) g.p_halo_dirty[1] = true;
    if (id == X_RTEMP) g.rt_halo_dirty[1] = true;
    HIP_OK_R(hipStreamSynchronize(g.stream), -1);
    return 0;
}

// =============================================================================
// Learned warm start (no reference counterpart; tea_warmstart.h)
// =============================================================================

EXPORT int tea_engine_warm_start_load(const float *params, int n)
{
    if (!g.initialised) { die("%s called before initialise_cuda_", __func__); return -1; }
    if (params && n != ws::NPARAM) {
        die("tea_engine_warm_start_load: expected %d parameters, got %d", ws::NPARAM, n);
        return -1;
    }
    if (params && !ws::shape_ok(g.ny, g.nx)) {
        die("tea_engine_warm_start_load: the network needs x_cells %% 64 == 0 and y_cells %% 4 == 0 (%d x %d)", g.nx,
            g.ny);
        return -1;
    }
    if (!g.ws_slab) {
        const int hist = read_int("tl_warm_start_history");
        if (hist != -1 && (hist < 1 || hist > WS_HMAX)) {
            die("tl_warm_start_history %d outside 1..%d", hist, WS_HMAX);
            return -1;
        }
        g.ws_hist = hist == -1 ? 8 : hist;
        const std::string norm = read_string("tl_warm_start_norm");
        if (!norm.empty() && norm != "residual" && norm != "energy") {
            die("tl_warm_start_norm %s: expected residual or energy", norm.c_str());
            return -1;
        }
        g.ws_minres = norm != "energy";
        const size_t elems = (size_t)g.pitch * (size_t)g.rows + 16;
        HIP_OK_R(hipMalloc(&g.ws_slab, (size_t)(g.ws_hist + 1) * elems * sizeof(double)), -1);
        HIP_OK_R(hipMemset(g.ws_slab, 0, (size_t)(g.ws_hist + 1) * elems * sizeof(double)), -1);
        for (int j = 0; j <= g.ws_hist; j++) g.ws_prev[j] = g.ws_slab + (g.fs[F_U] - g.f[F_U]) + (size_t)j * elems;
    }
    if (params) {
        if (!g.ws_net.params) {
            HIP_OK_R(ws::net_alloc(g.ws_net), -1);
            HIP_OK_R(ws::acts_alloc(g.ws_acts, 1, g.ny, g.nx, false), -1);
            HIP_OK_R(hipMalloc(&g.ws_planes, 3 * (size_t)g.nx * g.ny * sizeof(float)), -1);
            const size_t elems = (size_t)g.pitch * (size_t)g.rows + 16;
            HIP_OK_R(hipMalloc(&g.ws_qa, 2 * elems * sizeof(double)), -1);
        }
        HIP_OK_R(ws::net_upload(g.ws_net, params, g.stream), -1);
    }
    g.ws_on = true;
    g.ws_use_net = params != nullptr;
    g.ws_nprev = 0;
    g.ws_dn = 0;
    return 0;
}

EXPORT void tea_engine_warm_start_reset(void)
{
    g.ws_nprev = 0;
    g.ws_dn = 0;
}

namespace {
// Solves the k x k system G c = b (G symmetric positive semi-definite) by
// Cholesky, dropping a direction whose pivot falls below 1e-12 of its
// diagonal (linearly dependent on the earlier ones); c = 0 for dropped ones.
constexpr int WS_DMAX = WS_HMAX + 1;
void ws_solve(int k, double G[WS_DMAX][WS_DMAX], const double *b, double *c)
{
    double L[WS_DMAX][WS_DMAX] = {{0}};
    bool use[WS_DMAX] = {false};
    for (int j = 0; j < k; j++) {
        double dj = G[j][j];
        for (int m = 0; m < j; m++) dj -= L[j][m] * L[j][m];
        if (!(G[j][j] > 0.0) || !(dj > 1e-12 * G[j][j])) continue;
        use[j] = true;
        L[j][j] = std::sqrt(dj);
        for (int i = j + 1; i < k; i++) {
            double v = G[i][j];
            for (int m = 0; m < j; m++) v -= L[i][m] * L[j][m];
            L[i][j] = v / L[j][j];
        }
    }
    double y[WS_DMAX] = {0};
    for (int i = 0; i < k; i++) {
        if (!use[i]) continue;
        double v = b[i];
        for (int m = 0; m < i; m++) v -= L[i][m] * y[m];
        y[i] = v / L[i][i];
    }
    for (int i = k - 1; i >= 0; i--) {
        c[i] = 0.0;
        if (!use[i]) continue;
        double v = y[i];
        for (int m = i + 1; m < k; m++) v -= L[m][i] * c[m];
        c[i] = v / L[i][i];
    }
}
}  // namespace

namespace {
// At most WS_GRAM_GRID workgroups, each over many tiles: the workgroup's
// reduction of up to 3 (HN + 1) + 5 values then costs little beside its cells
// (measured against 4096 and 16384 at 4096^2).
constexpr int WS_GRAM_GRID = 1024;
template <int HN> bool ws_gram_launch(const WsDirs &d)
{
    auto go = [&](auto kernel) {
        return launch_capped("ws_gram", WS_GRAM_GRID, kernel, interior(), g.pitch, d, (const double *)F(X_KX),
                             (const double *)F(X_KY), red_ctx());
    };
    if (d.net) return d.has_last ? go(k_ws_gram<HN, true, true>) : go(k_ws_gram<HN, false, true>);
    return d.has_last ? go(k_ws_gram<HN, true, false>) : go(k_ws_gram<HN, false, false>);
}
template <int HN> bool ws_apply_launch(const WsDirs &d, double *keep, const WsCoefs &cf, int apply)
{
    return launch("ws_apply", k_ws_apply<HN>, interior(), g.pitch, F(F_U), d, keep, cf, apply);
}
}  // namespace

EXPORT void tea_engine_warm_start_apply(int *applied)
{
    CHECK_INIT();
    *applied = 0;
    if (!g.ws_on) { die("tea_engine_warm_start_apply: tea_engine_warm_start_load was not called"); return; }
    const Ext in = interior();
    const int HN = g.ws_hist;
    WsDirs d{};
    d.u0 = (const double *)F(X_U0);
    for (int j = 0; j <= HN; j++) d.up[j] = g.ws_prev[j];
    d.net = nullptr;
    d.scale = 0.0;
    d.h = g.h;
    d.nx = g.nx;
    d.ny = g.ny;
    d.nv = g.ws_nprev < HN + 1 ? g.ws_nprev : HN + 1;   // differences held
    d.nh = d.nv < HN ? d.nv : HN;                        // directions in the guess
    d.has_last = d.nh > 0 && d.nh == d.nv;               // v_nh not held yet: A v_{nh-1} by the stencil
    d.la = d.nh > 1 ? d.up[d.nh - 2] : d.u0;
    d.lb = d.nh > 0 ? d.up[d.nh - 1] : d.u0;
    WsCoefs cf{};
    int apply = 0;
    if (g.ws_nprev >= 1) {
        if (g.ws_use_net) {
            if (!launch("ws_stats", k_ws_stats, in, g.pitch, d.u0, (const double *)g.ws_prev[0], red_ctx())) return;
            double sums[2];
            if (!fetch(sums, 2)) return;
            const double cells = (double)g.nx * (double)g.ny;
            const double sc = std::sqrt(sums[0] / cells), su = std::sqrt(sums[1] / cells);
            if (sc > 0.0 && su > 0.0) {
                const long n = (long)g.nx * g.ny;
                if (!launch("ws_features", k_ws_features, in, g.pitch, g.h, g.nx, n, (const double *)F(F_DENSITY),
                            d.u0, (const double *)g.ws_prev[0], 1.0 / sc, 1.0 / su, g.ws_planes))
                    return;
                {
                    Timed t("ws_network");
                    ws::forward(g.ws_net, g.ws_acts, g.ws_planes, ws::NFEAT, g.stream);
                }
                HIP_OK(hipGetLastError());
                // the network predicts (du - (u0 - up1)) / s: direction s * net
                d.net = g.ws_acts.out;
                d.scale = sc;
                const size_t elems = (size_t)g.pitch * (size_t)g.rows + 16;
                double *q = g.ws_qa + (g.fs[F_U] - g.f[F_U]), *aq = q + elems;
                if (!launch("ws_netdir", k_ws_netdir, in, g.pitch, d, (const double *)F(X_KX), (const double *)F(X_KY),
                            q, aq))
                    return;
                d.q = q;
                d.aq = aq;
            }
        }
        bool ok;
        switch (HN) {
        case 1: ok = ws_gram_launch<1>(d); break;
        case 2: ok = ws_gram_launch<2>(d); break;
        case 3: ok = ws_gram_launch<3>(d); break;
        case 4: ok = ws_gram_launch<4>(d); break;
        case 5: ok = ws_gram_launch<5>(d); break;
        case 6: ok = ws_gram_launch<6>(d); break;
        case 7: ok = ws_gram_launch<7>(d); break;
        default: ok = ws_gram_launch<8>(d); break;
        }
        if (!ok) return;
        const int NV = HN + 1, nv = d.nv;
        double r[MAX_RED] = {0};
        const int nred = d.net ? 3 * NV + 5 : (d.has_last ? 2 * NV + 1 : NV);
        if (!fetch(r, nred)) return;
        // the old differences v_i are this step's v_{i+1}: their dot products carry over
        if (g.ws_dn < nv - 1) { die("tea_engine_warm_start_apply: history bookkeeping lost"); return; }
        double Dn[WS_HMAX + 1][WS_HMAX + 1] = {{0}};
        for (int x = 0; x + 1 < nv; x++)
            for (int y = 0; y + 1 < nv; y++) Dn[x + 1][y + 1] = g.ws_D[x][y];
        for (int j = 0; j < nv; j++) Dn[0][j] = Dn[j][0] = r[j];
        for (int x = 0; x < nv; x++)
            for (int y = 0; y < nv; y++) g.ws_D[x][y] = Dn[x][y];
        g.ws_dn = nv;
        const double LL = r[NV], *X = r + NV + 1;                  // last . last, last . v_j
        const double AqAq = r[2 * NV + 1], qAq = r[2 * NV + 2], qv0 = r[2 * NV + 3], Aqlast = r[2 * NV + 4];
        const double *Y = r + 2 * NV + 5;                          // A q . v_j
        // A v_a = v_{a+1} (index a+1) or, for the last direction with has_last, the stencil's (-1)
        auto img = [&](int a) { return a + 1 < nv ? a + 1 : -1; };
        auto AA = [&](int a, int b) {   // A v_a . A v_b
            const int ia = img(a), ib = img(b);
            if (ia >= 0 && ib >= 0) return Dn[ia][ib];
            if (ia < 0 && ib < 0) return LL;
            return X[ia >= 0 ? ia : ib];
        };
        auto VA = [&](int a, int b) {   // v_a . A v_b
            const int ib = img(b);
            return ib >= 0 ? Dn[a][ib] : X[a];
        };
        double G[WS_DMAX][WS_DMAX] = {{0}}, b[WS_DMAX] = {0};
        for (int x = 0; x < d.nh; x++) {
            for (int y = x; y < d.nh; y++) G[x][y] = G[y][x] = g.ws_minres ? AA(x, y) : VA(x, y);
            b[x] = g.ws_minres ? (img(x) >= 0 ? Dn[img(x)][0] : X[0]) : Dn[x][0];
        }
        if (d.net) {
            for (int x = 0; x < d.nh; x++)
                G[x][HN] = G[HN][x] = g.ws_minres ? (img(x) >= 0 ? Y[img(x)] : Aqlast) : Y[x];
            G[HN][HN] = g.ws_minres ? AqAq : qAq;
            b[HN] = g.ws_minres ? Y[0] : qv0;
        }
        // active directions in order: the history differences held, [the network]
        int idx[WS_DMAX], k = 0;
        for (int j = 0; j < d.nh; j++) idx[k++] = j;
        if (d.net) idx[k++] = HN;
        double Gs[WS_DMAX][WS_DMAX], bs[WS_DMAX], cs[WS_DMAX];
        for (int a = 0; a < k; a++) {
            bs[a] = b[idx[a]];
            for (int q = 0; q < k; q++) Gs[a][q] = G[idx[a]][idx[q]];
        }
        ws_solve(k, Gs, bs, cs);
        for (int a = 0; a < k; a++) cf.c[idx[a]] = cs[a];
        apply = 1;
    }
    for (int a = 0; a <= WS_HMAX; a++) g.ws_coef[a] = 0.0;
    for (int a = 0; a <= HN; a++) g.ws_coef[a] = cf.c[a];
    // u0 goes to the oldest buffer of the ring, which then becomes up_1
    double *keep = g.ws_prev[HN];
    bool ok;
    switch (HN) {
    case 1: ok = ws_apply_launch<1>(d, keep, cf, apply); break;
    case 2: ok = ws_apply_launch<2>(d, keep, cf, apply); break;
    case 3: ok = ws_apply_launch<3>(d, keep, cf, apply); break;
    case 4: ok = ws_apply_launch<4>(d, keep, cf, apply); break;
    case 5: ok = ws_apply_launch<5>(d, keep, cf, apply); break;
    case 6: ok = ws_apply_launch<6>(d, keep, cf, apply); break;
    case 7: ok = ws_apply_launch<7>(d, keep, cf, apply); break;
    default: ok = ws_apply_launch<8>(d, keep, cf, apply); break;
    }
    if (!ok) return;
    for (int j = HN; j > 0; j--) g.ws_prev[j] = g.ws_prev[j - 1];
    g.ws_prev[0] = keep;
    g.ws_nprev = g.ws_nprev < HN + 1 ? g.ws_nprev + 1 : HN + 1;
    *applied = apply;
}

// The coefficients of the last guess: c[0], c[1] the first two history
// differences, c[2] the network's direction.
EXPORT void tea_engine_warm_start_coefficients(double *c)
{
    c[0] = g.ws_coef[0];
    c[1] = g.ws_hist > 1 ? g.ws_coef[1] : 0.0;
    c[2] = g.ws_coef[g.ws_hist];
}

EXPORT int tea_engine_copy_field(int id, double *dst)
{
    if (g.initialised) settle_pending();   // rtemp and utemp owed by tea_engine_ppcg_begin_ur
    if (!g.initialised || id < 1 || id >= X_COUNT || !dst) return -1;
    const double *src = g.fs[id] + (long)g.h + (long)g.h * g.pitch;
    HIP_OK_R(hipMemcpy2DAsync(dst, (size_t)g.nx * sizeof(double), src, (size_t)g.pitch * sizeof(double),
                              (size_t)g.nx * sizeof(double), (size_t)g.ny, hipMemcpyDeviceToDevice, g.stream),
             -1);
    HIP_OK_R(hipStreamSynchronize(g.stream), -1);
    return 0;
}
