// tea_driver.cpp -- host driver of the TeaLeaf solver engine.
//
// Reads a tea.in deck with the semantics of parse.f90 / read_input.f90, sets the
// chunk up as start.f90 does (one chunk, all faces external), runs the time
// loop of diffuse.f90 and the solver control of tea_solve.f90 with the host
// numerics of tea_leaf_cheby.f90 / tea_leaf_ppcg.f90, and calls the backend's
// 29 kernel entry points in the same order as the Fortran driver.
//
// One deliberate difference: tea_solve.f90:224-247 runs the PPCG
// re-initialisation (ppcg_init steps 2 and 3 and the inner steps) for the
// Chebyshev solver as well; the CUDA layer registers those kernels for PPCG only
// (init_cuda.cu:262-275) and CUDALAUNCH looks them up with std::map::at
// (cuda_common.hpp:125-129), so the reference aborts there.  This driver runs
// that block for PPCG only, which is the only way the Chebyshev path completes.
#include "tea_driver.h"

#include <dlfcn.h>
#include <sys/time.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

// ---------------------------------------------------------------- constants
// data.f90:37-75
constexpr int CHUNK_LEFT = 1, CHUNK_RIGHT = 2, CHUNK_BOTTOM = 3, CHUNK_TOP = 4, EXTERNAL_FACE = -1;
constexpr int FIELD_DENSITY = 1, FIELD_ENERGY0 = 2, FIELD_ENERGY1 = 3, FIELD_U = 4, FIELD_P = 5,
              FIELD_SD = 6, NUM_FIELDS = 7;
constexpr int G_RECT = 1, G_CIRC = 2, G_POINT = 3;
constexpr int CONDUCTIVITY = 1, RECIP_CONDUCTIVITY = 2;
constexpr int TL_PREC_NONE = 1, TL_PREC_JAC_DIAG = 2, TL_PREC_JAC_BLOCK = 3;
// Default-kind (single precision) real literals of the Fortran sources keep
// their float value when they meet REAL(KIND=8) variables.
constexpr double F32(float x) { return (double)x; }
constexpr double G_SMALL = F32(1.0e-16f);  // data.f90:26
constexpr int G_IBIG = 640000;

struct Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] void fail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Fail(buf);
}

tea_step_callback g_step_cb = nullptr;
void *g_step_user = nullptr;

double wall_time()
{
    struct timeval t;
    gettimeofday(&t, nullptr);
    return (double)t.tv_sec + (double)t.tv_usec * 1.0e-6;
}

// ---------------------------------------------------------------- backend
struct Backend {
    void *handle = nullptr;
#define TEA_FN(name, ret, ...) ret (*name)(__VA_ARGS__) = nullptr;
#include "tea_abi_list.inc"
#undef TEA_FN
    // extensions
    void (*set_deck_path)(const char *) = nullptr;
    void (*set_error_mode)(int) = nullptr;
    int (*last_error)(char *, int) = nullptr;
    void (*clear_error)(void) = nullptr;
    void (*synchronize)(void) = nullptr;
    int (*set_precision)(int) = nullptr;
    int (*mp_leave)(void) = nullptr;
    void (*mp_setup)(void) = nullptr;
    void (*mp_begin)(double *) = nullptr;
    void (*mp_end)(void) = nullptr;
    void (*mp_replace)(double, double *, double *) = nullptr;
    int (*warm_start_load)(const float *, int) = nullptr;
    void (*warm_start_apply)(int *) = nullptr;
    void (*ppcg_begin)(int, double, double *) = nullptr;
    void (*ppcg_ur)(double, double, double *) = nullptr;
    void (*ppcg_end)(double *) = nullptr;
    void (*ppcg_inner_n)(int, int) = nullptr;
    void (*ppcg_begin_ur)(int, double, double, double *, double *) = nullptr;
    void (*ppcg_end_begin_ur)(double, double, double *, double *, double *) = nullptr;
    void (*ppcg_end_begin_ur_inner)(double, double, int, double *, double *, double *) = nullptr;
    void (*ppcg_iteration)(int, double, double, int, double *, double *, double *) = nullptr;
    void (*cheby_iterate_n)(int, int) = nullptr;
    int (*ppcg_inner_plan)(int) = nullptr;
    void (*cg_begin_ur)(int, double, double, double *, double *) = nullptr;

    void check() const
    {
        char msg[1024];
        if (last_error && last_error(msg, sizeof msg)) {
            if (clear_error) clear_error();
            fail("kernel layer: %s", msg);
        }
    }
};

Backend load_backend(const char *path)
{
    Backend b;
    b.handle = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!b.handle) fail("cannot load backend %s: %s", path, dlerror());
#define TEA_FN(name, ret, ...)                                                   \
    b.name = reinterpret_cast<ret (*)(__VA_ARGS__)>(dlsym(b.handle, #name));    \
    if (!b.name) fail("backend %s does not export %s", path, #name);
#include "tea_abi_list.inc"
#undef TEA_FN
    b.set_deck_path = reinterpret_cast<void (*)(const char *)>(dlsym(b.handle, "tea_engine_set_deck_path"));
    b.set_error_mode = reinterpret_cast<void (*)(int)>(dlsym(b.handle, "tea_engine_set_error_mode"));
    b.last_error = reinterpret_cast<int (*)(char *, int)>(dlsym(b.handle, "tea_engine_last_error"));
    b.clear_error = reinterpret_cast<void (*)(void)>(dlsym(b.handle, "tea_engine_clear_error"));
    b.synchronize = reinterpret_cast<void (*)(void)>(dlsym(b.handle, "tea_engine_synchronize"));
    b.set_precision = reinterpret_cast<int (*)(int)>(dlsym(b.handle, "tea_engine_set_precision"));
    b.mp_leave = reinterpret_cast<int (*)(void)>(dlsym(b.handle, "tea_engine_mp_leave"));
    b.mp_setup = reinterpret_cast<void (*)(void)>(dlsym(b.handle, "tea_engine_mp_setup"));
    b.mp_begin = reinterpret_cast<void (*)(double *)>(dlsym(b.handle, "tea_engine_mp_begin"));
    b.mp_end = reinterpret_cast<void (*)(void)>(dlsym(b.handle, "tea_engine_mp_end"));
    b.mp_replace = reinterpret_cast<void (*)(double, double *, double *)>(dlsym(b.handle, "tea_engine_mp_replace"));
    b.warm_start_load = reinterpret_cast<int (*)(const float *, int)>(dlsym(b.handle, "tea_engine_warm_start_load"));
    b.warm_start_apply = reinterpret_cast<void (*)(int *)>(dlsym(b.handle, "tea_engine_warm_start_apply"));
    b.ppcg_begin = reinterpret_cast<void (*)(int, double, double *)>(dlsym(b.handle, "tea_engine_ppcg_begin"));
    b.ppcg_ur = reinterpret_cast<void (*)(double, double, double *)>(dlsym(b.handle, "tea_engine_ppcg_ur"));
    b.ppcg_end = reinterpret_cast<void (*)(double *)>(dlsym(b.handle, "tea_engine_ppcg_end"));
    b.ppcg_inner_n = reinterpret_cast<void (*)(int, int)>(dlsym(b.handle, "tea_engine_ppcg_inner_n"));
    b.ppcg_begin_ur =
        reinterpret_cast<void (*)(int, double, double, double *, double *)>(dlsym(b.handle, "tea_engine_ppcg_begin_ur"));
    b.ppcg_inner_plan = reinterpret_cast<int (*)(int)>(dlsym(b.handle, "tea_engine_ppcg_inner_plan"));
    b.ppcg_end_begin_ur = reinterpret_cast<void (*)(double, double, double *, double *, double *)>(
        dlsym(b.handle, "tea_engine_ppcg_end_begin_ur"));
    b.ppcg_end_begin_ur_inner = reinterpret_cast<void (*)(double, double, int, double *, double *, double *)>(
        dlsym(b.handle, "tea_engine_ppcg_end_begin_ur_inner"));
    b.ppcg_iteration = reinterpret_cast<void (*)(int, double, double, int, double *, double *, double *)>(
        dlsym(b.handle, "tea_engine_ppcg_iteration"));
    b.cheby_iterate_n = reinterpret_cast<void (*)(int, int)>(dlsym(b.handle, "tea_engine_cheby_iterate_n"));
    b.cg_begin_ur =
        reinterpret_cast<void (*)(int, double, double, double *, double *)>(dlsym(b.handle, "tea_engine_cg_begin_ur"));
    return b;
}

// ---------------------------------------------------------------- deck
struct State {
    bool defined = false;
    double density = 0.0, energy = 0.0;
    double xmin = 0.0, xmax = 0.0, ymin = 0.0, ymax = 0.0, radius = 0.0;
    int geometry = 0;
};

struct Deck {
    double xmin = 0.0, ymin = 0.0, xmax = 100.0, ymax = 100.0;
    int x_cells = 10, y_cells = 10;
    double end_time = 10.0;
    int end_step = G_IBIG;
    int summary_frequency = 10;
    double dtinit = F32(0.1f);       // read_input.f90:49
    int max_iters = 1000;
    double eps = F32(1.0e-10f);      // read_input.f90:52
    int coefficient = CONDUCTIVITY;
    bool profiler_on = false;
    int tl_ch_cg_presteps = 25;
    double tl_ch_cg_epslim = 1.0;
    bool tl_check_result = false;
    int tl_preconditioner_type = TL_PREC_NONE;
    bool reflective_boundary = false;
    int tl_ppcg_inner_steps = -1;
    bool tl_use_chebyshev = false, tl_use_cg = false, tl_use_ppcg = false, tl_use_jacobi = true;
    bool verbose_on = false;
    int halo_exchange_depth = 1;
    int test_problem = 0;
    std::vector<State> states;  // 1-based in the deck, 0-based here
    // refinement on the FP64 true residual (no reference counterpart;
    // read_input.f90 ignores these keys)
    int refine_bits = 0;         // tl_inner_precision fp32|fp64; 0: the reference's solve
    bool restart = false;        // tl_mixed_refinement restart|reliable
    double reliable_delta = 0.01; // tl_reliable_delta: residual drop that triggers a replacement
    double inner_eps = 1.0e-4;   // tl_inner_eps: restart scheme's inner tolerance (relative)
    int max_refine = 30;         // tl_max_refinements: restart scheme's limit
    std::string warm_start;      // tl_warm_start <weight file>: learned initial guess
};

// The text after `key` on the deck line that starts with it, case kept (the
// keyword parser lower-cases and truncates lines; a file path must not be).
std::string raw_value(const std::string &path, const char *key)
{
    std::ifstream in(path);
    std::string l;
    const size_t kl = std::strlen(key);
    while (std::getline(in, l)) {
        size_t a = l.find_first_not_of(" \t");
        if (a == std::string::npos || l.size() - a < kl) continue;
        std::string head = l.substr(a, kl);
        for (auto &ch : head) ch = (char)std::tolower((unsigned char)ch);
        if (head != key) continue;
        size_t b = a + kl;
        while (b < l.size() && (l[b] == ' ' || l[b] == '\t' || l[b] == '=')) b++;
        size_t e = b;
        while (e < l.size() && l[e] != ' ' && l[e] != '\t' && l[e] != '\r') e++;
        return l.substr(b, e - b);
    }
    return std::string();
}

// One logical deck line as parse_getline returns it (parse.f90:107-236):
// control characters blanked, text after '!' or ';' dropped, lower-cased,
// '=' and ',' turned into blanks; only lines inside *tea ... *endtea kept.
std::vector<std::vector<std::string>> deck_lines(const std::string &path)
{
    std::ifstream in(path);
    if (!in) fail("cannot open input deck %s", path.c_str());
    std::vector<std::vector<std::string>> out;
    std::string l;
    bool inside = false, joining = false;
    while (std::getline(in, l)) {
        if (l.size() > 100) l.resize(100);  // READ (a100)
        for (auto &ch : l)
            if ((unsigned char)ch < 32 || (unsigned char)ch > 128) ch = ' ';
        auto cut = l.find('!');
        if (cut != std::string::npos) l.resize(cut);
        cut = l.find(';');
        if (cut != std::string::npos) l.resize(cut);
        for (auto &ch : l) ch = (char)std::tolower((unsigned char)ch);
        for (auto &ch : l)
            if (ch == '=' || ch == ',') ch = ' ';
        std::istringstream ss(l);
        std::vector<std::string> words;
        std::string w;
        while (ss >> w) words.push_back(w);
        if (words.empty()) continue;
        if (words[0][0] == '*') {
            if (words[0] == "*tea") inside = true;
            else if (words[0] == "*endtea") inside = false;
            continue;
        }
        if (!inside) continue;
        // a lone backslash continues the logical line on the next line (parse.f90:260)
        const bool cont = words.back() == "\\";
        if (cont) words.pop_back();
        if (joining) out.back().insert(out.back().end(), words.begin(), words.end());
        else out.push_back(words);
        joining = cont;
    }
    return out;
}

// parse_getival: READ (I7), i.e. the first 7 characters.
int getival(const std::string &w)
{
    std::string s = w.substr(0, 7);
    char *end = nullptr;
    long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || *end != 0) fail("Error attempting to convert to integer:%s", w.c_str());
    return (int)v;
}

// parse_getrval: READ (E27.20).
double getrval(const std::string &w)
{
    std::string s = w.substr(0, 27);
    char *end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end != 0) fail("Error attempting to convert to real:%s", w.c_str());
    return v;
}

// read_input.f90:24-329
Deck read_deck(const std::string &path)
{
    Deck d;
    auto lines = deck_lines(path);
    int state_max = 0;
    for (auto &ws : lines)
        for (size_t k = 0; k < ws.size(); k++)
            if (ws[k] == "state") {
                if (k + 1 >= ws.size()) fail("state without a number");
                state_max = std::max(state_max, getival(ws[k + 1]));
                break;
            }
    if (state_max < 1) fail("read_input: No states defined.");
    d.states.assign(state_max, State());

    for (auto &ws : lines) {
        size_t k = 0;
        auto next = [&](const char *key) -> const std::string & {
            if (k + 1 >= ws.size()) fail("missing value after %s", key);
            return ws[++k];
        };
        for (; k < ws.size(); k++) {
            const std::string &w = ws[k];
            if (w == "initial_timestep") d.dtinit = getrval(next("initial_timestep"));
            else if (w == "end_time") d.end_time = getrval(next("end_time"));
            else if (w == "end_step") d.end_step = getival(next("end_step"));
            else if (w == "xmin") d.xmin = getrval(next("xmin"));
            else if (w == "xmax") d.xmax = getrval(next("xmax"));
            else if (w == "ymin") d.ymin = getrval(next("ymin"));
            else if (w == "ymax") d.ymax = getrval(next("ymax"));
            else if (w == "x_cells") d.x_cells = getival(next("x_cells"));
            else if (w == "y_cells") d.y_cells = getival(next("y_cells"));
            else if (w == "visit_frequency") (void)getival(next("visit_frequency"));
            else if (w == "summary_frequency") d.summary_frequency = getival(next("summary_frequency"));
            else if (w == "tl_ch_cg_presteps") d.tl_ch_cg_presteps = getival(next("tl_ch_cg_presteps"));
            else if (w == "tl_ppcg_inner_steps") d.tl_ppcg_inner_steps = getival(next("tl_ppcg_inner_steps"));
            else if (w == "tl_ch_cg_epslim") d.tl_ch_cg_epslim = getrval(next("tl_ch_cg_epslim"));
            else if (w == "tl_check_result") d.tl_check_result = true;
            else if (w == "tl_preconditioner_type") {
                for (k++; k < ws.size(); k++) {
                    if (ws[k] == "none") d.tl_preconditioner_type = TL_PREC_NONE;
                    else if (ws[k] == "jac_diag") d.tl_preconditioner_type = TL_PREC_JAC_DIAG;
                    else if (ws[k] == "jac_block") d.tl_preconditioner_type = TL_PREC_JAC_BLOCK;
                    else fail("read_input: Invalid preconditioner type specified: %s", ws[k].c_str());
                }
            } else if (w == "use_cuda_kernels") {
            } else if (w == "verbose_on") d.verbose_on = true;
            else if (w == "tl_use_jacobi") { d.tl_use_chebyshev = d.tl_use_cg = d.tl_use_ppcg = false; d.tl_use_jacobi = true; }
            else if (w == "tl_use_cg") { d.tl_use_chebyshev = d.tl_use_ppcg = d.tl_use_jacobi = false; d.tl_use_cg = true; }
            else if (w == "tl_use_ppcg") { d.tl_use_chebyshev = d.tl_use_cg = d.tl_use_jacobi = false; d.tl_use_ppcg = true; }
            else if (w == "tl_use_chebyshev") { d.tl_use_cg = d.tl_use_ppcg = d.tl_use_jacobi = false; d.tl_use_chebyshev = true; }
            else if (w == "reflective_boundary") d.reflective_boundary = true;
            else if (w == "profiler_on") d.profiler_on = true;
            else if (w == "halo_depth") {
                d.halo_exchange_depth = getival(next("halo_depth"));
                if (d.halo_exchange_depth < 1) fail("read_input: Invalid halo exchange depth specified");
            } else if (w == "tl_max_iters") d.max_iters = getival(next("tl_max_iters"));
            else if (w == "tl_eps") d.eps = getrval(next("tl_eps"));
            else if (w == "tl_coefficient_density") d.coefficient = CONDUCTIVITY;
            else if (w == "tl_coefficient_inverse_density") d.coefficient = RECIP_CONDUCTIVITY;
            else if (w == "test_problem") d.test_problem = getival(next("test_problem"));
            else if (w == "tl_inner_precision") {
                const std::string &p = next("tl_inner_precision");
                if (p == "fp32") d.refine_bits = 32;
                else if (p == "fp64") d.refine_bits = 64;
                else fail("read_input: tl_inner_precision must be fp32 or fp64, not %s", p.c_str());
            } else if (w == "tl_mixed_refinement") {
                const std::string &p = next("tl_mixed_refinement");
                if (p == "restart") d.restart = true;
                else if (p == "reliable") d.restart = false;
                else fail("read_input: tl_mixed_refinement must be reliable or restart, not %s", p.c_str());
            } else if (w == "tl_reliable_delta") d.reliable_delta = getrval(next("tl_reliable_delta"));
            else if (w == "tl_warm_start") {
                next("tl_warm_start");
                d.warm_start = raw_value(path, "tl_warm_start");
            }
            else if (w == "tl_inner_eps") d.inner_eps = getrval(next("tl_inner_eps"));
            else if (w == "tl_max_refinements") d.max_refine = getival(next("tl_max_refinements"));
            else if (w == "state") {
                int s = getival(next("state"));
                if (s < 1 || s > state_max) fail("read_input: bad state number %d", s);
                State &st = d.states[s - 1];
                if (st.defined) fail("read_input: State defined twice.");
                st.defined = true;
                for (k++; k < ws.size(); k++) {
                    const std::string &sw = ws[k];
                    if (sw == "xmin") st.xmin = getrval(next("xmin"));
                    else if (sw == "ymin") st.ymin = getrval(next("ymin"));
                    else if (sw == "xmax") st.xmax = getrval(next("xmax"));
                    else if (sw == "ymax") st.ymax = getrval(next("ymax"));
                    else if (sw == "radius") st.radius = getrval(next("radius"));
                    else if (sw == "density") st.density = getrval(next("density"));
                    else if (sw == "energy") st.energy = getrval(next("energy"));
                    else if (sw == "geometry") {
                        const std::string &gw = next("geometry");
                        if (gw == "rectangle") st.geometry = G_RECT;
                        else if (gw == "circle") st.geometry = G_CIRC;
                        else if (gw == "point") st.geometry = G_POINT;
                    }
                }
            }
            // unknown keywords are ignored (read_input.f90:131-289 has no default)
        }
    }

    if (d.tl_ppcg_inner_steps == -1) {
        // 4*INT(SQRT(SQRT(REAL(x_cells*y_cells)))) in default (single) precision
        float c = (float)(d.x_cells * d.y_cells);
        d.tl_ppcg_inner_steps = 4 * (int)std::sqrt(std::sqrt(c));
    }
    if (d.refine_bits && !(d.tl_use_cg || d.tl_use_ppcg || d.tl_use_chebyshev))
        fail("read_input: tl_inner_precision needs tl_use_cg, tl_use_ppcg or tl_use_chebyshev");
    if (d.refine_bits && d.reflective_boundary)
        fail("read_input: tl_inner_precision does not support reflective_boundary");
    if (!d.warm_start.empty() && !(d.tl_use_cg || d.tl_use_ppcg))
        fail("read_input: tl_warm_start needs tl_use_cg or tl_use_ppcg");
    if (!d.warm_start.empty() && d.reflective_boundary)
        fail("read_input: tl_warm_start does not support reflective_boundary");
    if (d.refine_bits && !(d.reliable_delta > 0.0 && d.reliable_delta < 1.0))
        fail("read_input: tl_reliable_delta must lie in (0, 1)");
    if (d.halo_exchange_depth > 1 && d.tl_preconditioner_type == TL_PREC_JAC_BLOCK && d.tl_use_ppcg)
        fail("read_input: Unable to use nonstandard halo depth with block jacobi preconditioner");

    // state extents pulled in by 1/100 of a cell (read_input.f90:312-327)
    const double dx = (d.xmax - d.xmin) / (double)(float)d.x_cells;
    const double dy = (d.ymax - d.ymin) / (double)(float)d.y_cells;
    for (size_t n = 1; n < d.states.size(); n++) {
        d.states[n].xmin = d.states[n].xmin + (dx / 100.0);
        d.states[n].ymin = d.states[n].ymin + (dy / 100.0);
        d.states[n].xmax = d.states[n].xmax - (dx / 100.0);
        d.states[n].ymax = d.states[n].ymax - (dy / 100.0);
    }
    return d;
}

// ---------------------------------------------------------------- host numerics
// tqli of tea_leaf_cheby.f90:46-96 (QL with implicit shifts on a symmetric
// tridiagonal matrix; d and e are 1-based, e sub-diagonal shifted as eoshift).
int tqli(std::vector<double> &d, std::vector<double> &e, int n)
{
    // e(:) = eoshift(e(:), 1)
    for (int i = 1; i < n; i++) e[i] = e[i + 1];
    e[n] = 0.0;
    for (int l = 1; l <= n; l++) {
        int iter = 0;
        for (;;) {
            int m;
            for (m = l; m <= n - 1; m++) {
                double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) + dd == dd) break;
            }
            if (m == l) break;
            if (iter == 30) return 1;
            iter++;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::sqrt(g * g + 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(std::fabs(r), g));
            double s = 1.0, c = 1.0, p = 0.0;
            bool restart = false;
            for (int i = m - 1; i >= l; i--) {
                double f = s * e[i];
                double b = c * e[i];
                r = std::sqrt(f * f + g * g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] = d[i + 1] - p;
                    e[m] = 0.0;
                    restart = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
            }
            if (restart) continue;
            d[l] = d[l] - p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return 0;
}

// tea_calc_eigenvalues, tea_leaf_cheby.f90:98-145 (Lanczos tridiagonal from CG alpha/beta).
int calc_eigenvalues(const std::vector<double> &cg_alphas, const std::vector<double> &cg_betas,
                     double &eigmin, double &eigmax, int nsteps)
{
    std::vector<double> diag(nsteps + 2, 0.0), offdiag(nsteps + 2, 0.0);
    for (int n = 1; n <= nsteps; n++) {
        diag[n] = 1.0 / cg_alphas[n];
        if (n > 1) diag[n] = diag[n] + cg_betas[n - 1] / cg_alphas[n - 1];
        if (n < nsteps) offdiag[n + 1] = std::sqrt(cg_betas[n]) / cg_alphas[n];
    }
    int info = tqli(diag, offdiag, nsteps);
    if (info != 0) return info;
    std::sort(diag.begin() + 1, diag.begin() + 1 + nsteps);
    eigmin = diag[1];
    eigmax = diag[nsteps];
    if (eigmin < 0.0 || eigmax < 0.0) info = 1;
    return info;
}

// tea_calc_ch_coefs, tea_leaf_cheby.f90:147-174 (1-based arrays).
void calc_ch_coefs(std::vector<double> &ch_alphas, std::vector<double> &ch_betas, double eigmin,
                   double eigmax, double &theta, int n_coefs)
{
    theta = (eigmax + eigmin) / 2;
    const double delta = (eigmax - eigmin) / 2;
    const double sigma = theta / delta;
    double rho_old = 1.0 / sigma;
    if ((int)ch_alphas.size() < n_coefs + 1) {
        ch_alphas.resize(n_coefs + 1);
        ch_betas.resize(n_coefs + 1);
    }
    for (int n = 1; n <= n_coefs; n++) {
        const double rho_new = 1.0 / (2.0 * sigma - rho_old);
        ch_alphas[n] = rho_new * rho_old;
        ch_betas[n] = 2.0 * rho_new / delta;
        rho_old = rho_new;
    }
}

// Fortran Ew.d output (e.g. E26.17): 0.ddddE+xx right-justified.
std::string fortran_e(double x, int w, int d)
{
    char buf[128];
    if (x == 0.0) {
        std::string s = "0." + std::string(d, '0') + "E+00";
        if ((int)s.size() < w) s = std::string(w - s.size(), ' ') + s;
        return s;
    }
    snprintf(buf, sizeof buf, "%.*e", d - 1, x);  // d.ddd..e+xx
    std::string s(buf);
    bool neg = s[0] == '-';
    if (neg) s = s.substr(1);
    auto epos = s.find('e');
    std::string mant = s.substr(0, epos);
    int ex = std::atoi(s.c_str() + epos + 1);
    std::string digits = mant.substr(0, 1) + mant.substr(2);
    snprintf(buf, sizeof buf, "%s0.%sE%c%02d", neg ? "-" : "", digits.c_str(), (ex + 1) < 0 ? '-' : '+',
             std::abs(ex + 1));
    s = buf;
    if ((int)s.size() < w) s = std::string(w - s.size(), ' ') + s;
    return s;
}

// ---------------------------------------------------------------- the run
class Run {
public:
    Run(const Deck &deck, Backend &be, FILE *log, tea_run_result *res) : d(deck), b(be), out(log), R(res) {}

    void start();
    void diffuse();

private:
    const Deck &d;
    Backend &b;
    FILE *out;
    tea_run_result *R;

    int chunk_neighbours[4] = {EXTERNAL_FACE, EXTERNAL_FACE, EXTERNAL_FACE, EXTERNAL_FACE};
    int step = 0;
    double time = 0.0, dt = 0.0;
    bool complete = false;
    double rx = 0.0, ry = 0.0;

    // SAVEd across tea_leaf calls (tea_solve.f90:55,58)
    double eigmin = 0.0, eigmax = 0.0, theta = 0.0, cn = 0.0;
    bool first = true;

    void update_halo(const int *fields, int depth);
    void field_summary();
    void load_warm_start();
    void tea_leaf();
    struct Solve {
        int n;                  // loop index at exit (iteration count + 1)
        double error;           // the solver's own error measure at exit
        double initial;         // sqrt of the initial residual the loop was given
        int ppcg_inner;         // PPCG inner steps taken
        // error is the square root of the FP64 2-norm of u0 - A u for the final
        // u, computed by calc_residual and calc_2norm (or the fused replacement,
        // the same expressions) after u last changed: the refinement loops
        bool exact = false;
    };
    Solve iterate(double initial_residual_sq, double eps);
    Solve refine(double initial_residual_sq, double target_residual_sq, int *refinements, int *iterations,
                 int *inner);
    Solve refine_reliable(double initial_residual_sq, double target_residual_sq, int *updates, int *iterations,
                          int *inner);

    // residual replacement state of refine_reliable (off in the reference's solve)
    struct Reliable {
        bool on = false;
        int bits = 64;          // precision of the iteration fields
        double target = 0.0;    // stop when ||u0 - A u||_2 < target
        double rr_ref = 0.0;    // ||r||^2 (FP64, true) at the last replacement
        double q_ref = 0.0;     // the solver's rrn at the last replacement
        bool have_q = false;
        bool fp32_live = false; // the backend is in FP32 mode with u32 pending
        bool converged = false;
        bool stagnated = false; // the true residual stopped falling: FP64's attainable accuracy
        double q_last = 0.0;    // the last two measures passed to reliable_step
        double q_prev = 0.0;
        int nq = 0;
        int stalls = 0;
        int updates = 0;
        double rr_final = 0.0;
    } rel;
    bool reliable_step(double q, bool refresh_z, bool *replaced);
    bool replacement_unlikely() const;
    void run_ppcg_inner_steps(int *ppcg_inner_iters, bool fused = false);
    // the fused outer iteration (tea_engine_ppcg_begin/ur/end) is used when the
    // backend has it and no halo update separates the fused calls
    bool fused_outer() const
    {
        const char *env = std::getenv("TEA_FUSED_OUTER");
        return b.ppcg_begin && b.ppcg_ur && b.ppcg_end && !d.reflective_boundary && !(env && env[0] == '0');
    }
    void log(const char *fmt, ...);
};

void Run::log(const char *fmt, ...)
{
    if (!out) return;
    va_list ap;
    va_start(ap, fmt);
    vfprintf(out, fmt, ap);
    va_end(ap);
}

// update_halo.f90:30-85 at one chunk: no exchange; reflective update only when asked.
void Run::update_halo(const int *fields, int depth)
{
    if (d.reflective_boundary) {
        int f[NUM_FIELDS];
        std::memcpy(f, fields, sizeof f);
        int dep = depth;
        b.update_halo_kernel_cuda_(chunk_neighbours, f, &dep);
        b.check();
    }
}

// field_summary.f90:27-117
void Run::field_summary()
{
    double vol = 0.0, mass = 0.0, ie = 0.0, temp = 0.0;
    double tv = 0.0, tm = 0.0, ti = 0.0, tt = 0.0;
    b.field_summary_kernel_cuda_(&tv, &tm, &ti, &tt);
    b.check();
    vol = vol + tv;
    mass = mass + tm;
    ie = ie + ti;
    temp = temp + tt;
    if (R->n_summaries < TEA_MAX_SUMMARIES) {
        int k = R->n_summaries++;
        R->summary_step[k] = step;
        R->summary[k][0] = time;
        R->summary[k][1] = vol;
        R->summary[k][2] = mass;
        R->summary[k][3] = mass / vol;
        R->summary[k][4] = ie;
        R->summary[k][5] = temp;
    }
    log("\n Time %.15E\n", time);
    log("%13s%26s%26s%26s%26s%26s\n", "", "Volume", "Mass", "Density", "Energy", "U");
    log(" step:%7d%s%s%s%s%s\n\n", step, fortran_e(vol, 26, 17).c_str(), fortran_e(mass, 26, 17).c_str(),
        fortran_e(mass / vol, 26, 17).c_str(), fortran_e(ie, 26, 17).c_str(), fortran_e(temp, 26, 17).c_str());
    if (complete && d.test_problem >= 1) {
        static const double qa[7] = {0.0, 157.55084183279294, 106.27221178646569, 99.955877498324000,
                                     97.277332050749976, 95.462351583362249, 95.174738768320850};
        if (d.test_problem <= 6) {
            double qa_diff = std::fabs((100.0 * (temp / qa[d.test_problem])) - 100.0);
            log("Test problem %4d is within%s%% of the expected solution\n", d.test_problem, fortran_e(qa_diff, 16, 7).c_str());
            log(qa_diff < 0.001 ? " This test is considered PASSED\n" : " This is test is considered NOT PASSED\n");
        }
    }
}

// start.f90:24-113
void Run::start()
{
    time = 0.0;
    step = 0;
    dt = d.dtinit;
    int x_min = 1, x_max = d.x_cells, y_min = 1, y_max = d.y_cells;
    b.initialise_cuda_(&x_min, &x_max, &y_min, &y_max);
    b.check();

    // initialise_chunk.f90:31-44
    double dx = (d.xmax - d.xmin) / (double)(float)d.x_cells;
    double dy = (d.ymax - d.ymin) / (double)(float)d.y_cells;
    double xmin = d.xmin + dx * (double)(float)(1 - 1);
    double ymin = d.ymin + dy * (double)(float)(1 - 1);
    b.initialise_chunk_kernel_cuda_(&xmin, &ymin, &dx, &dy);
    b.check();

    // generate_chunk.f90:27-61
    const int ns = (int)d.states.size();
    std::vector<double> sd(ns), se(ns), sx0(ns), sx1(ns), sy0(ns), sy1(ns), srad(ns);
    std::vector<int> sgeo(ns);
    for (int s = 0; s < ns; s++) {
        sd[s] = d.states[s].density;
        se[s] = d.states[s].energy;
        sx0[s] = d.states[s].xmin;
        sx1[s] = d.states[s].xmax;
        sy0[s] = d.states[s].ymin;
        sy1[s] = d.states[s].ymax;
        srad[s] = d.states[s].radius;
        sgeo[s] = d.states[s].geometry;
    }
    int g_rect = G_RECT, g_circ = G_CIRC, g_point = G_POINT;
    b.generate_chunk_kernel_cuda_(&ns, sd.data(), se.data(), sx0.data(), sx1.data(), sy0.data(), sy1.data(),
                                  srad.data(), sgeo.data(), &g_rect, &g_circ, &g_point);
    b.check();

    int fields[NUM_FIELDS] = {0};
    fields[FIELD_DENSITY - 1] = 1;
    fields[FIELD_ENERGY0 - 1] = 1;
    fields[FIELD_ENERGY1 - 1] = 1;
    update_halo(fields, d.halo_exchange_depth);

    b.set_field_kernel_cuda_();
    b.check();
    field_summary();

    if (!d.warm_start.empty()) load_warm_start();
}

// Weight file: "TLWS0001", int32 count, count float32 (little-endian); the
// value "extrapolate" selects the same energy-minimising guess without a
// network (previous-step differences only).
void Run::load_warm_start()
{
    if (!b.warm_start_load || !b.warm_start_apply) fail("backend has no warm-start entry points (tl_warm_start)");
    if (d.warm_start == "extrapolate") {
        if (b.warm_start_load(nullptr, 0) != 0) b.check();
        b.check();
        log("Warm start from previous steps (no network)\n");
        return;
    }
    std::ifstream in(d.warm_start, std::ios::binary);
    if (!in) fail("tl_warm_start: cannot open %s", d.warm_start.c_str());
    char magic[8];
    int32_t n = 0;
    in.read(magic, 8);
    in.read(reinterpret_cast<char *>(&n), 4);
    if (!in || std::memcmp(magic, "TLWS0001", 8) != 0 || n <= 0 || n > (1 << 24))
        fail("tl_warm_start: %s is not a weight file", d.warm_start.c_str());
    std::vector<float> params(n);
    in.read(reinterpret_cast<char *>(params.data()), (std::streamsize)n * 4);
    if (!in) fail("tl_warm_start: %s is truncated", d.warm_start.c_str());
    if (b.warm_start_load(params.data(), n) != 0) b.check();
    b.check();
    log("Warm start network %s (%d parameters)\n", d.warm_start.c_str(), n);
}

// tea_leaf_ppcg.f90:174-230
// fused: sd/rtemp/utemp were set by tea_engine_ppcg_ur and z = utemp comes
// with tea_engine_ppcg_end (no reflective halo update to interleave)
void Run::run_ppcg_inner_steps(int *ppcg_inner_iters, bool fused)
{
    int fields[NUM_FIELDS] = {0};
    fields[FIELD_U - 1] = 1;
    update_halo(fields, 1);
    if (d.tl_ppcg_inner_steps < 0) return;
    double th = theta;
    if (!fused) {
        b.tea_leaf_ppcg_init_sd_new_kernel_cuda_(&th);
        b.check();
    }
    const char *env = std::getenv("TEA_INNER_PASSES");
    if (fused_outer() && b.ppcg_inner_n && b.ppcg_inner_plan && !(env && env[0] == '0')) {
        // one chunk, no reflective update: the steps run on the interior in
        // order, several per pass as the library plans them
        for (int step = 1; step <= d.tl_ppcg_inner_steps;) {
            const int n = b.ppcg_inner_plan(d.tl_ppcg_inner_steps - step + 1);
            if (n < 1) fail("tea_engine_ppcg_inner_plan returned no steps");
            b.ppcg_inner_n(step, n);
            b.check();
            step += n;
        }
        if (!fused) {
            b.tea_leaf_ppcg_update_z_kernel_cuda_();
            b.check();
        }
        return;
    }
    const int h = d.halo_exchange_depth;
    for (int cur = 1; cur <= d.tl_ppcg_inner_steps; cur += h) {
        int f[NUM_FIELDS] = {0};
        f[FIELD_SD - 1] = 1;
        update_halo(f, h);
        for (int bounds_extra = h - 1; bounds_extra >= 0; bounds_extra--) {
            int inner_step = cur + h - 1 - bounds_extra;
            int be = bounds_extra;
            b.tea_leaf_ppcg_inner_kernel_cuda_(&inner_step, &be, chunk_neighbours);
            b.check();
            update_halo(f, 1);  // update_boundary(fields, 1)
            if (inner_step == d.tl_ppcg_inner_steps) break;
        }
    }
    if (!fused) {
        b.tea_leaf_ppcg_update_z_kernel_cuda_();
        b.check();
    }
    (void)ppcg_inner_iters;
}

// tea_solve.f90:37-509
void Run::tea_leaf()
{
    double exact_error = 0.0, initial_residual = 0.0;
    int fields[NUM_FIELDS];

    if (d.coefficient != RECIP_CONDUCTIVITY && d.coefficient != CONDUCTIVITY)
        fail("tea_leaf: unknown coefficient option");

    std::fill(fields, fields + NUM_FIELDS, 0);
    fields[FIELD_ENERGY1 - 1] = 1;
    fields[FIELD_DENSITY - 1] = 1;
    update_halo(fields, d.halo_exchange_depth);

    // tea_leaf_common.f90:10-45: one chunk, one tile, every face external
    {
        int zero_boundary[4] = {1, 1, 1, 1};
        int reflective = d.reflective_boundary ? 1 : 0;
        int coef = d.coefficient;
        double dtv = dt;
        b.tea_leaf_common_init_kernel_cuda_(&coef, &dtv, &rx, &ry, zero_boundary, &reflective);
        b.check();
    }

    std::fill(fields, fields + NUM_FIELDS, 0);
    fields[FIELD_U - 1] = 1;
    update_halo(fields, 1);

    // With a warm start the target stays tl_eps times the residual of the
    // u = u^n guess; the solve starts from the network's guess.
    double target_residual = -1.0;
    int warm = 0;
    if (!d.warm_start.empty()) {
        b.tea_leaf_calc_residual_cuda_();
        b.check();
        int which = 1;
        b.tea_leaf_calc_2norm_kernel_cuda_(&which, &target_residual);
        b.check();
        b.warm_start_apply(&warm);
        b.check();
        update_halo(fields, 1);
    }
    b.tea_leaf_calc_residual_cuda_();
    b.check();
    {
        int which = 1;
        b.tea_leaf_calc_2norm_kernel_cuda_(&which, &initial_residual);
        b.check();
    }
    if (target_residual < 0.0) target_residual = initial_residual;
    if (warm)
        log("Warm start: residual %s of the u^n guess\n",
            fortran_e(std::sqrt(std::fabs(initial_residual)) / std::sqrt(std::fabs(target_residual)), 14, 7).c_str());
    int refinements = 0, iterations = 0, ppcg_inner_iters = 0;
    Solve sol;
    if (d.refine_bits && (d.restart || d.tl_use_chebyshev)) {
        // Chebyshev has no recurrence residual to replace: it always refines by restarts
        sol = refine(initial_residual, target_residual, &refinements, &iterations, &ppcg_inner_iters);
    } else if (d.refine_bits) {
        sol = refine_reliable(initial_residual, target_residual, &refinements, &iterations, &ppcg_inner_iters);
    } else {
        sol = iterate(target_residual, d.eps);
        iterations = sol.n - 1;
        ppcg_inner_iters = sol.ppcg_inner;
    }
    const double error = sol.error;
    initial_residual = sol.initial;

    if (d.tl_check_result && sol.exact && !d.reflective_boundary) {
        // the refinement loop's last true residual is this one: the same
        // kernels on the same u (no halo update changes u's halo here)
        exact_error = error;
    } else if (d.tl_check_result) {
        std::fill(fields, fields + NUM_FIELDS, 0);
        fields[FIELD_U - 1] = 1;
        update_halo(fields, 1);
        b.tea_leaf_calc_residual_cuda_();
        b.check();
        int which = 1;
        b.tea_leaf_calc_2norm_kernel_cuda_(&which, &exact_error);
        b.check();
        exact_error = std::sqrt(exact_error);
    }

    log("Conduction error %s\n", fortran_e(error / initial_residual, 14, 7).c_str());
    if (d.tl_check_result) log("EXACT error calculated as%s\n", fortran_e(exact_error / initial_residual, 14, 7).c_str());
    log("Iteration count %8d\n", iterations);
    if (d.tl_use_ppcg) log("PPCG Iteration count%8d (Total %8d)\n", ppcg_inner_iters, ppcg_inner_iters + iterations);
    if (d.refine_bits) log("FP%d refinements %8d\n", d.refine_bits, refinements);

    const int k = R->n_steps;
    if (k < TEA_MAX_STEPS) {
        R->iterations[k] = iterations;
        R->refinements[k] = refinements;
        R->warm_start[k] = warm;
        R->ppcg_inner[k] = ppcg_inner_iters;
        R->conduction_error[k] = error / initial_residual;
        R->exact_error[k] = d.tl_check_result ? exact_error / initial_residual : -1.0;
        R->initial_residual[k] = initial_residual;
    }
    R->eigmin = eigmin;
    R->eigmax = eigmax;

    b.tea_leaf_common_finalise_kernel_cuda_();
    b.check();
    std::fill(fields, fields + NUM_FIELDS, 0);
    fields[FIELD_ENERGY1 - 1] = 1;
    update_halo(fields, 1);
}

// The solver loop of tea_solve.f90:100-419 (after the initial residual): CG
// pre-steps, the switch to Chebyshev or PPCG, and the iterations to
// |error| < eps * sqrt(initial_residual_sq).
Run::Solve Run::iterate(double initial_residual_sq, double eps)
{
    int n;
    double old_error, error = 0.0, initial_residual = initial_residual_sq;
    double rro = 0.0, pw = 0.0, rrn = 0.0, alpha, beta;
    const int max_iters = d.max_iters;
    std::vector<double> cg_alphas(max_iters + 2, 0.0), cg_betas(max_iters + 2, 0.0);
    std::vector<double> ch_alphas(max_iters + 2, 0.0), ch_betas(max_iters + 2, 0.0);
    int est_itc = 0, cheby_calc_steps = 0, max_cheby_iters = 0, ppcg_inner_iters = 0;
    bool ch_switch_check = false, converged = false;
    const bool fused = d.tl_use_ppcg && fused_outer();
    bool p_pending = false;       // fused: ppcg_calc_p(beta_pending) still to apply
    double beta_pending = 0.0;
    // fused, residual replacement on: the end of an outer iteration (update_z,
    // calc_rrn) is issued with the next iteration's begin, beta = rrn / rro
    // formed on the device (the recurrence's error is not a stopping test
    // here, and nothing is printed per iteration)
    const bool defer_end = fused && rel.on && !d.verbose_on && b.ppcg_begin_ur && b.ppcg_end_begin_ur;
    bool end_pending = false;
    double rro_begin = 0.0;       // r.z of the latest begin
    // deferred end: the next inner steps may be queued with it, before its
    // values are known, when no residual replacement is expected (a replacement
    // after all overwrites what they wrote, and they run again)
    const char *spec_env = std::getenv("TEA_SPECULATE");
    const char *passes_env = std::getenv("TEA_INNER_PASSES");
    const bool may_queue_inner = defer_end && b.ppcg_end_begin_ur_inner && b.ppcg_inner_n && b.ppcg_inner_plan &&
                                 d.tl_ppcg_inner_steps > 0 && !(passes_env && passes_env[0] == '0') &&
                                 !(spec_env && spec_env[0] == '0');
    const bool queue_always = spec_env && spec_env[0] == '2';   // every time it may (tests the undo)
    // Chebyshev iterations between norm checks in streaming passes
    // (tea_engine_cheby_iterate_n) unless TEA_CHEBY_PASSES=0
    const char *cheby_env = std::getenv("TEA_CHEBY_PASSES");
    const bool cheby_batch = d.tl_use_chebyshev && b.cheby_iterate_n && !d.reflective_boundary &&
                             !(cheby_env && cheby_env[0] == '0');
    // end not deferred and no residual replacement: the whole outer iteration
    // (begin_ur, the inner passes, end) in one call with one host wait
    const char *iter_env = std::getenv("TEA_PPCG_ITERATION");
    const bool whole_iteration = fused && !defer_end && !rel.on && b.ppcg_iteration && b.ppcg_inner_n &&
                                 b.ppcg_inner_plan && d.tl_ppcg_inner_steps >= 0 &&
                                 !(passes_env && passes_env[0] == '0') && !(iter_env && iter_env[0] == '0');
    int fields[NUM_FIELDS];
    std::fill(fields, fields + NUM_FIELDS, 0);

    old_error = initial_residual;
    initial_residual = std::sqrt(std::fabs(initial_residual));
    if (d.verbose_on) log(" Initial residual %.16E\n", initial_residual);

    const bool cg_family = d.tl_use_cg || d.tl_use_chebyshev || d.tl_use_ppcg;
    if (cg_family) {
        b.tea_leaf_cg_init_kernel_cuda_(&rro);
        b.check();
        std::fill(fields, fields + NUM_FIELDS, 0);
        fields[FIELD_U - 1] = 1;
        fields[FIELD_P - 1] = 1;
        update_halo(fields, 1);
        std::fill(fields, fields + NUM_FIELDS, 0);
        fields[FIELD_P - 1] = 1;
    } else if (d.tl_use_jacobi) {
        std::fill(fields, fields + NUM_FIELDS, 0);
        fields[FIELD_U - 1] = 1;
    }

    // fused CG iterations (tea_engine_cg_begin_ur): the calc_p of each
    // iteration is applied by the next call, or by the separate entry point
    // when the CG branch is left (not under residual replacement, whose
    // updates fall between calc_ur and calc_p's successor)
    const char *fused_cg_env = std::getenv("TEA_FUSED_CG");
    const bool fused_cg = cg_family && b.cg_begin_ur && !d.reflective_boundary && !rel.on &&
                          !(fused_cg_env && fused_cg_env[0] == '0');
    bool cg_p_pending = false;
    double cg_beta_pending = 0.0;
    auto flush_cg_p = [&]() {
        if (!cg_p_pending) return;
        b.tea_leaf_cg_calc_p_kernel_cuda_(&cg_beta_pending);
        b.check();
        cg_p_pending = false;
    };

    for (n = 1; n <= max_iters; n++) {
        if (!ch_switch_check) {
            if (cheby_calc_steps > 0) ch_switch_check = true;
            else if (!first && d.tl_use_ppcg && n > 1) ch_switch_check = true;
            else if (std::fabs(old_error) <= d.tl_ch_cg_epslim && n >= d.tl_ch_cg_presteps) ch_switch_check = true;
            else ch_switch_check = false;
        }
        if ((d.tl_use_chebyshev || d.tl_use_ppcg) && ch_switch_check) flush_cg_p();

        if ((d.tl_use_chebyshev || d.tl_use_ppcg) && ch_switch_check) {
            if (cheby_calc_steps == 0) {
                max_cheby_iters = max_iters - n + 2;
                if (first) {
                    int info = calc_eigenvalues(cg_alphas, cg_betas, eigmin, eigmax, n - 1);
                    first = false;
                    if (info != 0) fail("tea_leaf: Error in calculating eigenvalues");
                    eigmin = eigmin * F32(0.95f);  // tea_solve.f90:186-187
                    eigmax = eigmax * F32(1.05f);
                }
                if (d.tl_use_chebyshev) {
                    calc_ch_coefs(ch_alphas, ch_betas, eigmin, eigmax, theta, max_cheby_iters);
                    std::fill(fields, fields + NUM_FIELDS, 0);
                    fields[FIELD_U - 1] = 1;
                } else if (d.tl_use_ppcg) {
                    calc_ch_coefs(ch_alphas, ch_betas, eigmin, eigmax, theta, d.tl_ppcg_inner_steps);
                    int ns = d.tl_ppcg_inner_steps;
                    b.tea_leaf_ppcg_init_constants_cuda_(ch_alphas.data() + 1, ch_betas.data() + 1, &ns);
                    b.check();
                }
                cn = eigmax / eigmin;
                log("Switching after %3d CG its, error %s\n", n, fortran_e(rro, 15, 7).c_str());
                log("Eigen min%s Eigen max%s Condition number%14.6f Error%s\n", fortran_e(eigmin, 14, 6).c_str(),
                    fortran_e(eigmax, 14, 6).c_str(), cn, fortran_e(old_error, 14, 6).c_str());

                if (d.tl_use_ppcg) {
                    int st = 2;
                    b.tea_leaf_ppcg_init_kernel_cuda_(&st, &rro);
                    b.check();
                    run_ppcg_inner_steps(&ppcg_inner_iters);
                    ppcg_inner_iters += d.tl_ppcg_inner_steps;
                    st = 3;
                    b.tea_leaf_ppcg_init_kernel_cuda_(&st, &rro);
                    b.check();
                    std::fill(fields, fields + NUM_FIELDS, 0);
                    fields[FIELD_P - 1] = 1;
                    update_halo(fields, 1);
                }
            }

            if (d.tl_use_chebyshev) {
                if (cheby_calc_steps == 0) {
                    // tea_leaf_cheby_first_step, tea_leaf_cheby.f90:176-220 (error argument is old_error)
                    double bb = 0.0;
                    int which = 0;
                    b.tea_leaf_calc_2norm_kernel_cuda_(&which, &bb);
                    b.check();
                    int nc = max_cheby_iters;
                    double th = theta;
                    b.tea_leaf_cheby_init_kernel_cuda_(ch_alphas.data() + 1, ch_betas.data() + 1, &nc, &th);
                    b.check();
                    update_halo(fields, 1);
                    int one = 1;
                    b.tea_leaf_cheby_iterate_kernel_cuda_(&one);
                    b.check();
                    which = 1;
                    b.tea_leaf_calc_2norm_kernel_cuda_(&which, &old_error);
                    b.check();
                    const double it_alpha = d.eps * bb / (4.0 * old_error);
                    const double gamm = (std::sqrt(cn) - 1.0) / (std::sqrt(cn) + 1.0);
                    est_itc = (int)std::lround(std::log(it_alpha) / (2.0 * std::log(gamm)));
                    log("    est itc\n%11d\n", est_itc);
                    cheby_calc_steps = 1;
                } else if (cheby_batch) {
                    // iterations n .. last in streaming passes: up to the next
                    // norm check, tl_max_iters, the last coefficient, or the
                    // iteration after which the loop's own test stops it (between
                    // checks error only has its square root taken, below); the
                    // skipped iterations' bookkeeping is replayed as the loop does it
                    int last = n;
                    double e = error;
                    while (last - n + 1 < 100 && !(last >= est_itc && last % 10 == 0) && last < max_iters &&
                           cheby_calc_steps + (last + 1 - n) <= max_cheby_iters) {
                        const double e_after = std::sqrt(std::fabs(e));
                        if (std::fabs(e_after) < eps * initial_residual) break;
                        e = e_after;
                        last++;
                    }
                    b.cheby_iterate_n(cheby_calc_steps, last - n + 1);
                    b.check();
                    for (; n < last; n++) {
                        cheby_calc_steps = cheby_calc_steps + 1;
                        update_halo(fields, 1);
                        error = std::sqrt(std::fabs(error));
                        if (d.verbose_on) log(" Residual %.16E\n", error);
                        old_error = error;
                    }
                    if (n >= est_itc && n % 10 == 0) {
                        int which = 1;
                        b.tea_leaf_calc_2norm_kernel_cuda_(&which, &error);
                        b.check();
                    }
                } else {
                    int st = cheby_calc_steps;
                    b.tea_leaf_cheby_iterate_kernel_cuda_(&st);
                    b.check();
                    if (n >= est_itc && n % 10 == 0) {
                        int which = 1;
                        b.tea_leaf_calc_2norm_kernel_cuda_(&which, &error);
                        b.check();
                    }
                }
            } else if (d.tl_use_ppcg && whole_iteration) {
                // the calls of the branch below in one: begin_ur, the inner
                // steps, end (rrn)
                double two[2], rrn_ur = 0.0;
                b.ppcg_iteration(p_pending ? 1 : 0, beta_pending, theta, d.tl_ppcg_inner_steps, two, &rrn_ur, &rrn);
                b.check();
                pw = two[0];
                rro = two[1];
                ppcg_inner_iters += d.tl_ppcg_inner_steps;
                beta = rrn / rro;
                p_pending = true;
                beta_pending = beta;
                error = rrn;
                rro = rrn;
            } else if (d.tl_use_ppcg && fused) {
                // the same calls as below, fused; ppcg_calc_p of the previous
                // iteration is folded into this iteration's first pass
                double two[2];
                bool inner_queued = false;
                if (end_pending) {   // the previous iteration's end, then begin with beta formed on the device
                    double rrn_end = 0.0;
                    if (may_queue_inner && (queue_always || replacement_unlikely())) {
                        b.ppcg_end_begin_ur_inner(rro_begin, theta, d.tl_ppcg_inner_steps, &rrn_end, two, &rrn);
                        inner_queued = true;
                    } else {
                        b.ppcg_end_begin_ur(rro_begin, theta, &rrn_end, two, &rrn);
                    }
                    b.check();
                    beta = rrn_end / rro_begin;
                    end_pending = false;
                    p_pending = false;
                    pw = two[0];
                    rro = two[1];
                } else if (b.ppcg_begin_ur) {   // alpha = rro / pw formed on the device
                    b.ppcg_begin_ur(p_pending ? 1 : 0, beta_pending, theta, two, &rrn);
                    b.check();
                    p_pending = false;
                    pw = two[0];
                    rro = two[1];
                } else {
                    b.ppcg_begin(p_pending ? 1 : 0, beta_pending, two);
                    b.check();
                    p_pending = false;
                    pw = two[0];
                    rro = two[1];
                    alpha = rro / pw;
                    b.ppcg_ur(alpha, theta, &rrn);
                    b.check();
                }
                rro_begin = rro;
                bool replaced = false;
                const int updates_before = rel.updates;
                if (rel.on && reliable_step(rrn, true, &replaced)) {
                    converged = true;
                } else {
                    if (replaced) {  // sd, rtemp, utemp from the replaced r and z
                        double th = theta;
                        b.tea_leaf_ppcg_init_sd_new_kernel_cuda_(&th);
                        b.check();
                    }
                    if (!inner_queued || rel.updates != updates_before) run_ppcg_inner_steps(&ppcg_inner_iters, true);
                    ppcg_inner_iters += d.tl_ppcg_inner_steps;
                    if (defer_end) {
                        end_pending = true;
                    } else {
                        b.ppcg_end(&rrn);
                        b.check();
                        beta = rrn / rro;
                        p_pending = true;
                        beta_pending = beta;
                    }
                }
                error = rrn;
                rro = rrn;
            } else if (d.tl_use_ppcg) {
                b.tea_leaf_cg_calc_w_kernel_cuda_(&pw);
                b.check();
                b.tea_leaf_ppcg_store_r_kernel_cuda_();
                b.check();
                b.tea_leaf_ppcg_calc_2norm_kernel_cuda_(&rro);
                b.check();
                alpha = rro / pw;
                b.tea_leaf_cg_calc_ur_kernel_cuda_(&alpha, &rrn);
                b.check();
                if (rel.on && reliable_step(rrn, true, nullptr)) {
                    converged = true;
                } else {
                    run_ppcg_inner_steps(&ppcg_inner_iters);
                    ppcg_inner_iters += d.tl_ppcg_inner_steps;
                    b.tea_leaf_ppcg_calc_rrn_kernel_cuda_(&rrn);
                    b.check();
                    beta = rrn / rro;
                    b.tea_leaf_ppcg_calc_p_kernel_cuda_(&beta);
                    b.check();
                }
                error = rrn;
                rro = rrn;
            }
            cheby_calc_steps = cheby_calc_steps + 1;
        } else if (cg_family && fused_cg) {
            // the calls below in one: [calc_p(beta) of the previous iteration], calc_w, calc_ur
            fields[FIELD_P - 1] = 1;
            b.cg_begin_ur(cg_p_pending ? 1 : 0, cg_beta_pending, rro, &pw, &rrn);
            b.check();
            alpha = rro / pw;
            cg_alphas[n] = alpha;
            beta = rrn / rro;
            cg_betas[n] = beta;
            cg_p_pending = true;
            cg_beta_pending = beta;
            error = rrn;
            rro = rrn;
        } else if (cg_family) {
            fields[FIELD_P - 1] = 1;
            b.tea_leaf_cg_calc_w_kernel_cuda_(&pw);
            b.check();
            alpha = rro / pw;
            cg_alphas[n] = alpha;
            b.tea_leaf_cg_calc_ur_kernel_cuda_(&alpha, &rrn);
            b.check();
            beta = rrn / rro;
            cg_betas[n] = beta;
            b.tea_leaf_cg_calc_p_kernel_cuda_(&beta);
            b.check();
            // a replaced r reaches z at the next cg_calc_ur
            if (rel.on && reliable_step(rrn, false, nullptr)) converged = true;
            error = rrn;
            rro = rrn;
        } else if (d.tl_use_jacobi) {
            b.tea_leaf_jacobi_solve_kernel_cuda_(&error);
            b.check();
        }

        update_halo(fields, 1);
        if (converged) break;

        error = std::sqrt(std::fabs(error));
        if (d.verbose_on) log(" Residual %.16E\n", error);
        if (std::fabs(error) < eps * initial_residual) break;
        old_error = error;
    }
    flush_cg_p();
    // a DO loop that runs to completion leaves n = max_iters + 1
    if (end_pending) {   // stopped by tl_max_iters: complete the last outer iteration
        b.ppcg_end(&rrn);
        b.check();
        error = std::sqrt(std::fabs(rrn));
    }

    return Solve{n, error, initial_residual, ppcg_inner_iters};
}

// Iterative refinement by restarts (no reference counterpart): the FP64
// residual r = u0 - A u is measured on the device; while
// ||r|| > tl_eps * ||r0|| the correction equation A d = r is solved by the
// same CG/PPCG/Chebyshev loop to tl_inner_eps relative, and u += d in FP64.
// With tl_inner_precision fp32 the inner solves run on FP32 fields (FP64 dot
// products and host numerics; d starts at 0); with fp64 they run on the FP64
// fields and update u in place.  The eigenvalue estimate comes from the first
// inner solve's CG pre-steps and is kept across solves.
Run::Solve Run::refine(double initial_residual_sq, double target_residual_sq, int *refinements, int *iterations,
                       int *inner)
{
    const bool fp32 = d.refine_bits == 32;
    if (fp32 && (!b.set_precision || !b.mp_setup || !b.mp_begin || !b.mp_end))
        fail("backend has no mixed-precision entry points (tl_inner_precision fp32)");
    const double ir = std::sqrt(std::fabs(target_residual_sq));
    if (fp32) {
        if (b.set_precision(32) != 0) b.check();
        b.set_precision(64);
        b.mp_setup();
        b.check();
    }
    double rr = initial_residual_sq;
    double err = std::sqrt(std::fabs(rr));
    *refinements = 0;
    *iterations = 0;
    *inner = 0;
    for (int k = 0; k < d.max_refine; k++) {
        if (k > 0) {
            b.tea_leaf_calc_residual_cuda_();
            b.check();
            int which = 1;
            b.tea_leaf_calc_2norm_kernel_cuda_(&which, &rr);
            b.check();
            err = std::sqrt(std::fabs(rr));
        }
        if (d.verbose_on) log(" Refinement %d true residual %s\n", k, fortran_e(err / ir, 14, 7).c_str());
        if (err < d.eps * ir) break;
        Solve s;
        if (fp32) {
            double rr32 = 0.0;
            b.set_precision(32);
            b.mp_begin(&rr32);
            b.check();
            s = iterate(rr32, d.inner_eps);
            b.mp_end();
            b.check();
            b.set_precision(64);
        } else {
            s = iterate(rr, d.inner_eps);
        }
        *refinements += 1;
        *iterations += s.n - 1;
        *inner += s.ppcg_inner;
    }
    if (*refinements == d.max_refine) {
        b.tea_leaf_calc_residual_cuda_();
        b.check();
        int which = 1;
        b.tea_leaf_calc_2norm_kernel_cuda_(&which, &rr);
        b.check();
        err = std::sqrt(std::fabs(rr));
    }
    return Solve{0, err, ir, *inner, true};
}

// Refinement with residual replacement ("reliable updates": Sleijpen and
// van der Vorst, Computing 56 (1996) 141-163; Clark et al., Comput. Phys.
// Commun. 181 (2010) 1517-1528), no reference counterpart.  The CG/PPCG loop
// runs on fields of `bits` precision (FP64 scalars and dot accumulators) on
// the correction equation A d = r, where u32 holds d since the last
// replacement.  Whenever the loop's own residual measure has dropped by
// tl_reliable_delta since the last replacement, or predicts that the target
// is met, the correction is folded into the FP64 u, the FP64 residual
// r = u0 - A u is measured, and the loop stops if ||r||_2 < tl_eps ||r0||_2;
// otherwise r replaces the loop's residual (in PPCG z = M^-1 r is recomputed
// at once; CG forms it at its next cg_calc_ur), keeping the search direction
// p, so the Krylov process is not restarted.
// At 64 bits the same loop is CG/PPCG with residual replacement and a stop
// test on the true residual.
Run::Solve Run::refine_reliable(double initial_residual_sq, double target_residual_sq, int *updates,
                                int *iterations, int *inner)
{
    const int bits = d.refine_bits;
    if (bits == 32 && (!b.set_precision || !b.mp_setup || !b.mp_begin || !b.mp_end))
        fail("backend has no mixed-precision entry points (tl_inner_precision fp32)");
    const double ir = std::sqrt(std::fabs(target_residual_sq));
    const double r0 = std::sqrt(std::fabs(initial_residual_sq));
    rel = Reliable();
    rel.on = true;
    rel.bits = bits;
    rel.target = d.eps * ir;
    rel.rr_ref = initial_residual_sq;
    rel.rr_final = initial_residual_sq;
    double rr = initial_residual_sq;
    Solve s{1, r0, ir, 0};
    if (!(r0 < rel.target) && r0 > 0.0) {
        if (bits == 32) {
            if (b.set_precision(32) != 0) b.check();
            b.mp_setup();
            b.check();
            b.mp_begin(&rr);
            b.check();
            rel.fp32_live = true;
        }
        s = iterate(rr, 0.0);
        if (rel.fp32_live) {  // stopped by tl_max_iters
            b.mp_end();
            b.check();
            b.set_precision(64);
            rel.fp32_live = false;
        }
        if (!rel.converged) {
            b.tea_leaf_calc_residual_cuda_();
            b.check();
            int which = 1;
            b.tea_leaf_calc_2norm_kernel_cuda_(&which, &rel.rr_final);
            b.check();
        }
    }
    rel.on = false;
    *updates = rel.updates;
    *iterations = s.n - 1;
    *inner = s.ppcg_inner;
    return Solve{s.n, std::sqrt(std::fabs(rel.rr_final)), ir, s.ppcg_inner, true};
}

// One residual-replacement test of refine_reliable, given the loop's own
// measure q: true when the solve ends (converged or stagnated).  After a
// replacement, *replaced (if given) says whether the caller still has to form
// sd, rtemp and utemp (tea_engine_mp_replace forms them itself).
// Whether the next reliable_step can be expected to replace nothing: no
// reference level yet, or the last two measures, extrapolated geometrically,
// keep the drop and the predicted residual a margin (1.5, or TEA_SPEC_MARGIN)
// above their thresholds (1.25 and 1.5 measured best on the 4096^2 bench
// deck, 0.5% ahead of 2).
bool Run::replacement_unlikely() const
{
    static const double margin = [] {
        const char *s = std::getenv("TEA_SPEC_MARGIN");
        const double v = s ? std::atof(s) : 0.0;
        return v >= 1.0 ? v : 1.5;
    }();
    if (!rel.on) return false;
    if (!rel.have_q) return true;
    if (rel.nq < 2 || !(std::fabs(rel.q_prev) > 0.0) || !(std::fabs(rel.q_ref) > 0.0)) return false;
    const double ratio = std::fabs(rel.q_last / rel.q_prev);
    const double q_next = std::fabs(rel.q_last) * (ratio < 1.0 ? ratio : 1.0);
    const double drop = std::sqrt(q_next / std::fabs(rel.q_ref));
    const double predicted = std::sqrt(std::fabs(rel.rr_ref)) * drop;
    return drop >= margin * d.reliable_delta && predicted >= margin * rel.target;
}

bool Run::reliable_step(double q, bool refresh_z, bool *replaced)
{
    if (replaced) *replaced = false;
    rel.q_prev = rel.q_last;
    rel.q_last = q;
    rel.nq++;
    if (!rel.have_q) {  // the first measure after a replacement is the reference level
        rel.q_ref = q;
        rel.have_q = true;
        return false;
    }
    const double drop = std::sqrt(std::fabs(q / rel.q_ref));
    const double predicted = std::sqrt(std::fabs(rel.rr_ref)) * drop;
    if (drop >= d.reliable_delta && predicted >= rel.target) return false;
    // FP32 PPCG: the replacement and the restart from it (r32, z, sd, rtemp,
    // utemp) in one engine call; when the solve ends here, the restart is
    // simply left unused
    const bool fused = rel.bits == 32 && refresh_z && replaced && b.mp_replace;
    double rr = 0.0;
    if (fused) {
        double rr32 = 0.0;
        b.mp_replace(theta, &rr, &rr32);
        b.check();
    } else {
        if (rel.bits == 32) {
            b.mp_end();
            b.check();
            b.set_precision(64);
            rel.fp32_live = false;
        }
        b.tea_leaf_calc_residual_cuda_();
        b.check();
        int which = 1;
        b.tea_leaf_calc_2norm_kernel_cuda_(&which, &rr);
        b.check();
    }
    auto leave_fp32 = [&] {   // the solve ends: its FP32 inner fields are not read again
        if (!fused) return;
        if (b.mp_leave) b.mp_leave();
        else b.set_precision(64);
        rel.fp32_live = false;
    };
    rel.updates++;
    rel.rr_final = rr;
    if (d.verbose_on) log(" Replacement %d true residual %.16E\n", rel.updates, std::sqrt(std::fabs(rr)));
    if (std::sqrt(std::fabs(rr)) < rel.target) {
        leave_fp32();
        rel.converged = true;
        return true;
    }
    // A replacement stalls when the true residual came out more than 4x above
    // what the loop's own measure predicted (the recurrence no longer tracks
    // it) or fell by less than 1% since the previous replacement.  Three in a
    // row mean the target lies below what FP64 resolves.
    const double now = std::sqrt(std::fabs(rr));
    const bool stall = now > 4.0 * predicted || now > 0.99 * std::sqrt(std::fabs(rel.rr_ref));
    rel.stalls = stall ? rel.stalls + 1 : 0;
    if (rel.stalls >= 3) {
        leave_fp32();
        rel.converged = true;
        rel.stagnated = true;
        log("True residual stagnated at %s of the target\n",
            fortran_e(std::sqrt(std::fabs(rr)) / rel.target, 14, 7).c_str());
        return true;
    }
    if (!fused) {
        if (rel.bits == 32) {
            b.set_precision(32);
            double rr32 = 0.0;
            b.mp_begin(&rr32);
            b.check();
            rel.fp32_live = true;
        }
        if (refresh_z) {
            int st = 2;
            double unused = 0.0;
            b.tea_leaf_ppcg_init_kernel_cuda_(&st, &unused);  // z = M^-1 r for the new r
            b.check();
        }
        if (replaced) *replaced = true;  // the caller forms sd, rtemp, utemp
    }
    rel.rr_ref = rr;
    rel.q_ref = q;
    return false;
}

// diffuse.f90:23-163
void Run::diffuse()
{
    const double timerstart = wall_time();
    for (;;) {
        const double step_time = wall_time();
        step = step + 1;
        dt = d.dtinit;  // timestep.f90 / calc_dt.f90: dt is fixed
        log(" Step %7d time %11.7f timestep  %9.2E\n", step, time, dt);
        tea_leaf();
        time = time + dt;
        if (d.summary_frequency != 0 && step % d.summary_frequency == 0) field_summary();
        if (b.synchronize) b.synchronize();
        const double now = wall_time();
        if (R->n_steps < TEA_MAX_STEPS) R->step_seconds[R->n_steps] = now - step_time;
        R->n_steps++;
        log(" Wall clock %.15g\n", now - timerstart);
        if (g_step_cb) g_step_cb(step, g_step_user);
        if (time + G_SMALL > d.end_time || step >= d.end_step) {
            complete = true;
            field_summary();
            log("\n Calculation complete\n Tea is finishing\n Wall clock %.15g\n", wall_time() - timerstart);
            break;
        }
    }
    R->total_seconds = wall_time() - timerstart;
}

}  // namespace

extern "C" void tea_driver_set_step_callback(tea_step_callback cb, void *user)
{
    g_step_cb = cb;
    g_step_user = user;
}

extern "C" int tea_driver_run(const char *deck_path, const char *backend_lib, const char *out_path,
                              tea_run_result *result)
{
    std::memset(result, 0, sizeof *result);
    FILE *log = nullptr;
    try {
        Deck deck = read_deck(deck_path);
        static std::string loaded_path;
        static Backend backend;
        if (!backend.handle || loaded_path != backend_lib) {
            backend = load_backend(backend_lib);
            loaded_path = backend_lib;
        }
        if (backend.set_error_mode) backend.set_error_mode(1);
        if (backend.clear_error) backend.clear_error();
        if (backend.set_deck_path) backend.set_deck_path(deck_path);
        if (out_path) {
            log = std::fopen(out_path, "w");
            if (!log) fail("cannot open %s", out_path);
        }
        Run run(deck, backend, log, result);
        run.start();
        run.diffuse();
    } catch (const std::exception &e) {
        result->error_code = 1;
        std::snprintf(result->error_msg, sizeof result->error_msg, "%s", e.what());
    }
    if (log) std::fclose(log);
    return result->error_code;
}
