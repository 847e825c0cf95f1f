// engine.cpp — C ABI over HIP + hipRTC (replaces U-CUDA/nvrtc_engine.cpp and the
// NVRTC/driver half of U-CUDA/parametric_engine.cpp).
//
// * A system (vars/params/rhs) is parsed once; its device functions are
//   generated by codegen and pasted, with the UcJob descriptor, in front of
//   the hand-written ensemble kernels (kernels/ensemble.hip, embedded at
//   build time). hipRTC compiles the program for gfx950 on first use.
// * Compiled modules are cached per (source, strict) with LRU eviction
//   (capacity 8), the pattern of NvrtcEngine's cache (nvrtc_engine.h:47-56).
// * The engine is internally locked; launches go on the caller's stream.
// * Axis tables live in small device buffers recycled once the event
//   recorded after the launch that used them has completed, so an enqueue
//   allocates only when every buffer is still in flight.
#include "../../include/ucuda.h"

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <list>
#include <map>
#include <tuple>
#include <unordered_map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "codegen.hpp"
#include "scratch.hpp"
#include "uc_job.h"

namespace {

#include "kernels_embed.inc"  // const char* kEnsembleSource, kJobHeader

constexpr int kMaxVars = 32;      // kMaxAmountOfX, U-CUDA/parametric_engine.cpp:30
constexpr int kMaxValues = 64;    // kMaxAmountOfValues, parametric_engine.cpp:31
constexpr int kBlock = 256;
constexpr size_t kModuleCacheSize = 8;

int cu_count(int device);

struct Module {
    std::string key;
    int pins = 0;  // live batches using the module: never evicted while pinned
    hipModule_t mod = nullptr;
    size_t code_size = 0;
    std::vector<char> code;
};

struct TableBuffer {
    double* dev = nullptr;
    double* host = nullptr;  // pinned staging, reused only after `done`
    size_t cap = 0;        // doubles
    hipEvent_t done = nullptr;
    bool in_flight = false;
};

uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ULL) {
    for (unsigned char c : s) { h ^= c; h *= 1099511628211ULL; }
    return h;
}

}  // namespace

struct uc_system {
    ucuda::System sys;
    bool custom = false;  // caller-supplied KRS step body instead of a right-hand side
    std::string device_source;
    int n_vars = 0;
    int n_values = 0;
    ucuda::OpCounts ops;
};

struct uc_engine {
    int device = -1;
    // One reference for the caller's handle plus one per live batch / server:
    // those hold module pins and table buffers of this engine, so the engine
    // is freed by whichever of uc_engine_destroy / uc_batch_destroy /
    // uc_server_destroy drops the last reference, in any order.
    std::atomic<int> refs{1};
    std::recursive_mutex mu;
    // `error` has a lock of its own, taken by nothing else, so it can be set
    // from any thread whatever other locks that thread holds
    mutable std::mutex err_mu;
    std::string error;
    // serialises hipRTC compiles, which run without `mu` (they take seconds)
    std::mutex compile_mu;
    std::list<Module> modules;      // front = most recently used
    std::vector<TableBuffer> tables;
    // uc_run_lle_noalloc: pinned host image of the tables (UC_SCRATCH_DOUBLES),
    // reused once `noalloc_done` (recorded after its copy) has completed;
    // both made when the engine is created, with the device's scratch address
    double* noalloc_host = nullptr;
    hipEvent_t noalloc_done = nullptr;
    std::string noalloc_error;      // why uc_run_lle_noalloc cannot run, if so

    int fail(int code, const std::string& msg) {
        std::lock_guard<std::mutex> lk(err_mu);
        error = msg;
        return code;
    }
    ~uc_engine() {
        for (auto& m : modules)
            if (m.mod) (void)hipModuleUnload(m.mod);
        for (auto& t : tables) {
            if (t.done) { (void)hipEventSynchronize(t.done); (void)hipEventDestroy(t.done); }
            if (t.dev) (void)hipFree(t.dev);
            if (t.host) (void)hipHostFree(t.host);
        }
        if (noalloc_done) { (void)hipEventSynchronize(noalloc_done); (void)hipEventDestroy(noalloc_done); }
        if (noalloc_host) (void)hipHostFree(noalloc_host);
    }
};

namespace {

std::string hip_msg(hipError_t e, const char* what) {
    return std::string(what) + ": " + hipGetErrorString(e);
}

std::string program_source(const uc_system* s, bool strict) {
    std::string src;
    src += "#define UC_STRICT ";
    src += strict ? "1\n" : "0\n";
    src += kJobHeader;
    src += "\n";
    src += s->device_source;
    src += "\n";
    src += kEnsembleSource;
    return src;
}

// ---- persistent code-object cache ----
// hipRTC takes seconds per module (every analysis family of a system is in
// one program), so compiled code objects are kept on disk, keyed by the
// program text, the compile options and the hipRTC version. Directory:
// $UC_CACHE_DIR (set but empty: no cache), else kcache/ beside this
// library. A file is "UCKC1 <second hash of the key> <size> <hash of the
// code>\n" + the code object; any mismatch (collision, truncation, damage)
// means a recompile. Writes go
// to a temporary file renamed into place, so concurrent processes never see
// a partial entry; any I/O failure just means compiling.
std::string kcache_dir() {
    const char* env = getenv("UC_CACHE_DIR");
    if (env) return std::string(env);
    Dl_info info;
    if (dladdr((void*)&kcache_dir, &info) && info.dli_fname) {
        const std::string p(info.dli_fname);
        const size_t slash = p.find_last_of('/');
        return (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/kcache";
    }
    return std::string();
}

std::string hex64(uint64_t v) {
    char b[17];
    snprintf(b, sizeof b, "%016llx", (unsigned long long)v);
    return std::string(b);
}

uint64_t fnv1a_bytes(const std::vector<char>& v) {
    uint64_t h = 1469598103934665603ULL;
    for (char c : v) { h ^= (unsigned char)c; h *= 1099511628211ULL; }
    return h;
}

std::string kcache_header(const std::string& tag, const std::vector<char>& code) {
    return tag + " " + hex64((uint64_t)code.size()) + " " + hex64(fnv1a_bytes(code)) + "\n";
}

// `tag` = "UCKC1 <second hash of the key>"; the rest of the header is
// checked against the code read back.
bool kcache_load(const std::string& path, const std::string& tag, std::vector<char>& code) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    std::string hdr(tag.size() + 35, '\0');  // tag + " " + 16 hex + " " + 16 hex + "\n"
    bool ok = fread(&hdr[0], 1, hdr.size(), f) == hdr.size() && hdr.compare(0, tag.size(), tag) == 0;
    if (ok) {
        char buf[1 << 16];
        size_t n;
        code.clear();
        while ((n = fread(buf, 1, sizeof buf, f)) > 0) code.insert(code.end(), buf, buf + n);
        ok = !ferror(f) && code.size() > 4 && memcmp(code.data(), "\x7f" "ELF", 4) == 0 &&
             kcache_header(tag, code) == hdr;
    }
    fclose(f);
    if (!ok) code.clear();
    return ok;
}

void kcache_store(const std::string& dir, const std::string& path, const std::string& tag,
                  const std::vector<char>& code) {
    (void)mkdir(dir.c_str(), 0755);  // may exist
    const std::string tmp = path + ".tmp." + std::to_string((long long)getpid()) + "." +
                            std::to_string((unsigned long long)std::hash<std::thread::id>()(std::this_thread::get_id()));
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) return;
    const std::string hdr = kcache_header(tag, code);
    const bool ok = fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size() &&
                    fwrite(code.data(), 1, code.size(), f) == code.size();
    if (fclose(f) == 0 && ok && rename(tmp.c_str(), path.c_str()) == 0) return;
    (void)unlink(tmp.c_str());
}

std::string module_key(const std::string& src, bool strict) {
    return std::to_string(fnv1a(src)) + (strict ? ":s" : ":f");
}

// The code object of a program: from the disk cache, else hipRTC for gfx950.
// Touches no engine state, so it runs without the engine lock.
// Fast modules fuse a * b + c inside one source expression only
// (-ffp-contract=on): which operations fuse is then fixed by the source
// text, not by how a kernel's code is laid out in basic blocks, so every
// kernel variant running the same cell code (direct, persistent, sliced, two
// cells per lane) gives the same bits. Strict modules fuse nothing.
int compile_code(const std::string& src, bool strict, std::vector<char>& code, std::string& err) {
    std::vector<const char*> opts = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                                     strict ? "-ffp-contract=off" : "-ffp-contract=on",
                                     "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    int vmaj = 0, vmin = 0;
    (void)hiprtcVersion(&vmaj, &vmin);
    std::string ident = src + "\n";
    for (const char* o : opts) ident += std::string(o) + " ";
    ident += "\nhiprtc " + std::to_string(vmaj) + "." + std::to_string(vmin);
    const std::string cdir = kcache_dir();
    const std::string cpath = cdir.empty() ? std::string() : cdir + "/" + hex64(fnv1a(ident)) + ".hsaco";
    const std::string ctag = "UCKC1 " + hex64(fnv1a(ident, 0x84222325cbf29ce4ULL));
    if (!cpath.empty() && kcache_load(cpath, ctag, code)) return UC_OK;
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "ucuda_ensemble.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        err = "hiprtcCreateProgram failed";
        return UC_ERR_COMPILE;
    }
    hiprtcResult rc = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
    if (rc != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        err = "hipRTC compile failed:\n" + log;
        return UC_ERR_COMPILE;
    }
    size_t size = 0;
    hiprtcGetCodeSize(prog, &size);
    code.resize(size);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    if (!cpath.empty()) kcache_store(cdir, cpath, ctag, code);
    return UC_OK;
}

Module* find_module(uc_engine* e, const std::string& key) {
    for (auto it = e->modules.begin(); it != e->modules.end(); ++it) {
        if (it->key == key) {
            e->modules.splice(e->modules.begin(), e->modules, it);
            return &e->modules.front();
        }
    }
    return nullptr;
}

// Inserts a compiled module at the front (under e->mu), with LRU eviction
// past the capacity, skipping modules pinned by live batches and servers.
Module* insert_module(uc_engine* e, const std::string& key, std::vector<char>&& code) {
    Module m;
    m.key = key;
    m.code = std::move(code);
    m.code_size = m.code.size();
    e->modules.push_front(std::move(m));
    size_t over = e->modules.size() > kModuleCacheSize ? e->modules.size() - kModuleCacheSize : 0;
    for (auto it = std::prev(e->modules.end()); over > 0 && it != e->modules.begin();) {
        auto cur = it--;
        if (cur->pins > 0) continue;
        if (cur->mod) (void)hipModuleUnload(cur->mod);
        e->modules.erase(cur);
        --over;
    }
    return &e->modules.front();
}

// Compiles the module of (s, strict) into the engine's cache WITHOUT holding
// e->mu, so a seconds-long hipRTC compile does not block the engine's other
// callers; entry points call it before taking the lock and then find the
// module in get_module. Failures are left for get_module to report.
void precompile(uc_engine* e, uc_system* s, bool strict) {
    if (!e || !s) return;
    const std::string src = program_source(s, strict);
    const std::string key = module_key(src, strict);
    {
        std::lock_guard<std::recursive_mutex> lk(e->mu);
        if (find_module(e, key)) return;
    }
    std::lock_guard<std::mutex> ck(e->compile_mu);
    {
        std::lock_guard<std::recursive_mutex> lk(e->mu);
        if (find_module(e, key)) return;  // compiled by another caller meanwhile
    }
    std::vector<char> code;
    std::string err;
    if (compile_code(src, strict, code, err) != UC_OK) return;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    if (!find_module(e, key)) insert_module(e, key, std::move(code));
}

// The module of (s, strict) from the cache (compiled here, under the lock,
// if precompile did not run or it was evicted since). With load=false only
// the code object is produced (works without a device). Caller holds e->mu.
int get_module(uc_engine* e, uc_system* s, bool strict, bool load, Module** out) {
    const std::string src = program_source(s, strict);
    const std::string key = module_key(src, strict);
    Module* m = find_module(e, key);
    if (!m) {
        std::vector<char> code;
        std::string err;
        const int rc = compile_code(src, strict, code, err);
        if (rc != UC_OK) return e->fail(rc, err);
        m = insert_module(e, key, std::move(code));
    }
    if (load && !m->mod) {
        hipError_t he = hipModuleLoadData(&m->mod, m->code.data());
        if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "hipModuleLoadData"));
    }
    *out = m;
    return UC_OK;
}

// Reference grid values (getValueByIdx / getValueByIdx_log,
// U-CUDA/cudaLibrary.cu:1344-1396), evaluated in double on the host.
void axis_values(const uc_axis& ax, double* out) {
    const int n = ax.n_pts;
    for (int i = 0; i < n; ++i) {
        if (ax.log_scale) {
            if (n == 1) { out[i] = std::pow(10.0, std::log10(ax.lo)); continue; }
            const double l0 = std::log10(ax.lo), l1 = std::log10(ax.hi);
            out[i] = std::pow(10.0, l0 + (double)i * ((l1 - l0) / (double)(n - 1)));
        } else {
            if (n == 1) { out[i] = ax.hi; continue; }
            const double t = (double)i / (double)(n - 1);
            out[i] = (1.0 - t) * ax.lo + t * ax.hi;
        }
    }
}

int validate_common(uc_engine* e, const uc_system* s, const uc_common& c, bool allow_step_axis) {
    if (!c.initial_conditions && c.ic_seed == 0) return e->fail(UC_ERR_ARGUMENT, "initial_conditions is NULL");
    if (!c.base_values && c.n_values > 0) return e->fail(UC_ERR_ARGUMENT, "base_values is NULL");
    if (c.n_values > kMaxValues) return e->fail(UC_ERR_ARGUMENT, "too many base_values");
    if (c.n_values < s->n_values) return e->fail(UC_ERR_ARGUMENT, "base_values shorter than the system's a[] (a[0] + params)");
    if (c.n_axes < 1 || c.n_axes > 2) return e->fail(UC_ERR_ARGUMENT, "n_axes must be 1 or 2");
    int step_axes = 0;
    for (int i = 0; i < c.n_axes; ++i) {
        const uc_axis& ax = c.axis[i];
        if (ax.n_pts <= 0) return e->fail(UC_ERR_ARGUMENT, "axis n_pts must be > 0");
        if (ax.log_scale && !(ax.lo > 0.0 && ax.hi > 0.0)) return e->fail(UC_ERR_ARGUMENT, "log scale needs lo > 0 and hi > 0");
        switch (ax.kind) {
        case UC_AXIS_PARAM:
            if (ax.index <= 0 || ax.index >= c.n_values || ax.index >= s->n_values)
                return e->fail(UC_ERR_ARGUMENT, "param_index out of range");
            break;
        case UC_AXIS_VAR:
            if (ax.index < 0 || ax.index >= s->n_vars) return e->fail(UC_ERR_ARGUMENT, "var_sweep_index out of range");
            break;
        case UC_AXIS_STEP:
            if (!allow_step_axis) return e->fail(UC_ERR_ARGUMENT, "this analysis cannot sweep h");
            ++step_axes;
            break;
        default: return e->fail(UC_ERR_ARGUMENT, "unknown axis kind");
        }
    }
    if (step_axes > 1) return e->fail(UC_ERR_ARGUMENT, "only one axis may sweep h");
    if (!(c.h > 0.0) && step_axes == 0) return e->fail(UC_ERR_ARGUMENT, "h must be > 0");
    if (!(c.t_max > 0.0)) return e->fail(UC_ERR_ARGUMENT, "t_max must be > 0");
    if (c.transient_time < 0.0) return e->fail(UC_ERR_ARGUMENT, "transient_time must be >= 0");
    if (c.precision != UC_PREC_F64 && c.precision != UC_PREC_F32C) return e->fail(UC_ERR_ARGUMENT, "unknown precision");
    if (c.cell_count < 0 || c.cell_begin < 0 || (c.cell_count == 0 && c.cell_begin != 0))
        return e->fail(UC_ERR_ARGUMENT, "cell range: need cell_count >= 0, cell_begin >= 0 (0 when cell_count is 0)");
    long long total = 1;
    for (int i = 0; i < c.n_axes; ++i) total *= c.axis[i].n_pts;
    if (c.cell_count > 0 && c.cell_begin + c.cell_count > total)
        return e->fail(UC_ERR_ARGUMENT, "cell range beyond the grid");
    return UC_OK;
}

// For entry points that run whole grids only.
int whole_grid_only(uc_engine* e, const uc_common& c, const char* what) {
    if (c.cell_count != 0) return e->fail(UC_ERR_ARGUMENT, std::string(what) + " runs whole grids (cell_count must be 0)");
    return UC_OK;
}

// Cells a request runs: its cell range, else the whole grid.
long long cells_of(const uc_common& c) {
    if (c.cell_count > 0) return c.cell_count;
    long long n = 1;
    for (int i = 0; i < c.n_axes; ++i) n *= c.axis[i].n_pts;
    return n;
}

size_t table_doubles(const uc_common& c) {
    const int nx = c.axis[0].n_pts, ny = c.n_axes > 1 ? c.axis[1].n_pts : 1;
    return (size_t)nx + (size_t)ny + (size_t)kMaxValues + (size_t)kMaxVars;
}

// Host image of a launch's input tables: axis values (nx, then ny), 64
// parameter slots, 32 initial-state slots.
void fill_tables(int n_vars, const uc_common& c, double* host) {
    const int nx = c.axis[0].n_pts, ny = c.n_axes > 1 ? c.axis[1].n_pts : 1;
    std::fill(host, host + table_doubles(c), 0.0);
    double* ax1 = host + nx;
    double* par = ax1 + ny;
    double* ic = par + kMaxValues;
    axis_values(c.axis[0], host);
    if (c.n_axes > 1) axis_values(c.axis[1], ax1);
    for (int i = 0; i < c.n_values; ++i) par[i] = c.base_values[i];
    for (int i = 0; i < n_vars; ++i) ic[i] = c.initial_conditions ? c.initial_conditions[i] : 0.0;
}

// The largest float <= v (FLT_MAX above its range): for a finite float x,
// x > float_at_most(v) exactly when (double)x > v.
float float_at_most(double v) {
    if (v >= (double)FLT_MAX) return FLT_MAX;
    float f = (float)v;
    if ((double)f > v) f = std::nextafter(f, -INFINITY);
    return f;
}

// Descriptor fields common to every analysis, pointing at device tables.
void wire_job(const uc_common& c, const double* dev, UcJob& j) {
    const int nx = c.axis[0].n_pts, ny = c.n_axes > 1 ? c.axis[1].n_pts : 1;
    std::memset(&j, 0, sizeof j);
    j.n_cells = c.cell_count > 0 ? c.cell_count : (long long)nx * ny;
    j.cell_offset = c.cell_count > 0 ? c.cell_begin : 0;
    j.stride = j.n_cells;
    j.nx = nx;
    j.ny = ny;
    for (int i = 0; i < 2; ++i) {
        j.axis_kind[i] = i < c.n_axes ? c.axis[i].kind : UC_AXIS_NONE;
        j.axis_index[i] = i < c.n_axes ? c.axis[i].index : 0;
    }
    j.axis_vals[0] = dev;
    j.axis_vals[1] = dev + nx;
    j.base_params = dev + nx + ny;
    j.base_ic = dev + nx + ny + kMaxValues;
    j.n_values = c.n_values;
    j.ic_seed = c.ic_seed;
    j.ic_lo = c.ic_lo;
    j.ic_hi = c.ic_hi;
    j.h = c.h;
    j.hh = 0.5 * c.h;
    j.h6 = c.h * (1.0 / 6.0);
    j.hf = (float)c.h;
    j.hhf = 0.5f * j.hf;
    j.h6f = j.hf * (1.0f / 6.0f);
    j.max_value = c.max_value;
    j.max_value_f = float_at_most(c.max_value);
    j.transient_time = c.transient_time;
    j.t_max = c.t_max;
    j.pre_scaller = 1;
}

// Uploads the tables into a recycled device buffer (pinned staging, reused
// once the event recorded after its launch has completed) and wires `j`.
int stage_tables(uc_engine* e, const uc_system* s, const uc_common& c, UcJob& j, hipStream_t stream, TableBuffer** used) {
    const size_t need = table_doubles(c);
    TableBuffer* tb = nullptr;
    for (auto& t : e->tables) {
        if (t.in_flight && hipEventQuery(t.done) == hipSuccess) t.in_flight = false;
        if (!t.in_flight && t.cap >= need) { tb = &t; break; }
    }
    if (!tb) {
        e->tables.emplace_back();
        tb = &e->tables.back();
        tb->cap = std::max<size_t>(need, 4096);
        hipError_t he = hipMalloc((void**)&tb->dev, tb->cap * sizeof(double));
        if (he == hipSuccess) he = hipHostMalloc((void**)&tb->host, tb->cap * sizeof(double), hipHostMallocDefault);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&tb->done, hipEventDisableTiming);
        if (he != hipSuccess) {
            if (tb->dev) (void)hipFree(tb->dev);
            if (tb->host) (void)hipHostFree(tb->host);
            e->tables.pop_back();
            return e->fail(UC_ERR_HIP, hip_msg(he, "table buffer allocation"));
        }
    }
    fill_tables(s->n_vars, c, tb->host);
    hipError_t he = hipMemcpyAsync(tb->dev, tb->host, need * sizeof(double), hipMemcpyHostToDevice, stream);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "hipMemcpyAsync(tables)"));
    wire_job(c, tb->dev, j);
    *used = tb;
    return UC_OK;
}

double step_of(const uc_common& c) {
    for (int i = 0; i < c.n_axes; ++i)
        if (c.axis[i].kind == UC_AXIS_STEP) return std::min(c.axis[i].lo, c.axis[i].hi);
    return c.h;
}

// Launches `name` over `lanes` lanes (default: one per cell).
int launch(uc_engine* e, Module* m, const char* name, UcJob& j, hipStream_t stream, TableBuffer* tb,
           long long lanes = -1) {
    hipFunction_t f;
    hipError_t he = hipModuleGetFunction(&f, m->mod, name);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, name));
    if (lanes < 0) lanes = j.n_cells;
    const long long blocks = (lanes + kBlock - 1) / kBlock;
    if (blocks > 0x7fffffffLL) return e->fail(UC_ERR_ARGUMENT, "grid too large");
    void* args[] = {&j};
    he = hipModuleLaunchKernel(f, (unsigned)blocks, 1, 1, kBlock, 1, 1, 0, stream, args, nullptr);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "hipModuleLaunchKernel"));
    if (!tb) return UC_OK;  // tables in the static scratch (uc_run_lle_noalloc)
    he = hipEventRecord(tb->done, stream);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "hipEventRecord"));
    tb->in_flight = true;
    return UC_OK;
}

int enter(uc_engine* e) {
    if (e->device < 0) return e->fail(UC_ERR_NO_DEVICE, "engine was created without a device");
    hipError_t he = hipSetDevice(e->device);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "hipSetDevice"));
    return UC_OK;
}

ucuda::System to_system(const uc_system_desc* d) {
    ucuda::System s;
    for (int i = 0; i < d->n_vars; ++i) s.vars.push_back(d->vars[i]);
    for (int i = 0; i < d->n_params; ++i) s.params.push_back(d->params[i]);
    for (int i = 0; i < d->n_vars; ++i) s.rhs.push_back(d->rhs[i]);
    s.latex = d->latex != 0;
    if (d->time_symbol) s.time_symbol = d->time_symbol;
    return s;
}

int copy_out(const std::string& text, char* buf, size_t cap, size_t* len) {
    if (len) *len = text.size() + 1;
    if (buf && cap > 0) {
        size_t n = std::min(cap - 1, text.size());
        std::memcpy(buf, text.data(), n);
        buf[n] = '\0';
    }
    return UC_OK;
}

}  // namespace

extern "C" {

const char* uc_version(void) { return "ucuda-mi355x 0.1"; }

int uc_engine_create(int device, uc_engine** out) {
    if (!out) return UC_ERR_ARGUMENT;
    auto e = std::make_unique<uc_engine>();
    e->device = device;
    if (device >= 0) {
        hipError_t he = hipSetDevice(device);
        if (he != hipSuccess) return UC_ERR_HIP;
        // set-up of uc_run_lle_noalloc (its failure leaves the other entry points usable)
        char msg[256] = {0};
        if (!uc_scratch_setup(device, msg, sizeof msg)) {
            e->noalloc_error = msg;
        } else {
            he = hipHostMalloc((void**)&e->noalloc_host, sizeof(double) * UC_SCRATCH_DOUBLES, hipHostMallocDefault);
            if (he == hipSuccess) he = hipEventCreateWithFlags(&e->noalloc_done, hipEventDisableTiming);
            if (he != hipSuccess) e->noalloc_error = hip_msg(he, "scratch staging set-up");
        }
    }
    *out = e.release();
    return UC_OK;
}

namespace {
void engine_release(uc_engine* e) {
    if (e && e->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) {
        if (e->device >= 0) (void)hipSetDevice(e->device);
        delete e;
    }
}
}  // namespace

void uc_engine_destroy(uc_engine* e) { engine_release(e); }

// A copy per calling thread: the engine's message may be rewritten by another
// thread at any time, the copy only by this thread's next call.
const char* uc_last_error(const uc_engine* e) {
    static thread_local std::string copy;
    if (!e) return "null engine";
    std::lock_guard<std::mutex> lk(e->err_mu);
    copy = e->error;
    return copy.c_str();
}

int uc_system_create(uc_engine* e, const uc_system_desc* desc, uc_system** out) {
    if (!e || !desc || !out) return UC_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    if (desc->n_vars <= 0 || desc->n_vars > kMaxVars) return e->fail(UC_ERR_ARGUMENT, "n_vars out of range");
    if (desc->n_params < 0 || desc->n_params + 1 > kMaxValues) return e->fail(UC_ERR_ARGUMENT, "n_params out of range");
    auto s = std::make_unique<uc_system>();
    try {
        s->sys = to_system(desc);
        s->device_source = ucuda::emit_system_source(s->sys, ucuda::Dialect::Hip);
        s->ops = ucuda::count_ops(s->sys);
    } catch (const std::exception& ex) {
        return e->fail(UC_ERR_PARSE, ex.what());
    }
    s->n_vars = desc->n_vars;
    s->n_values = desc->n_params + 1;
    *out = s.release();
    return UC_OK;
}

int uc_system_create_krs(uc_engine* e, const char* krs_body, int n_vars, int n_values, uc_system** out) {
    if (!e || !krs_body || !out) return UC_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    if (n_vars <= 0 || n_vars > kMaxVars) return e->fail(UC_ERR_ARGUMENT, "n_vars out of range");
    if (n_values < 1 || n_values > kMaxValues) return e->fail(UC_ERR_ARGUMENT, "n_values out of range");
    const std::string diag = ucuda::check_krs_indices(krs_body, n_vars, n_values);
    if (!diag.empty()) return e->fail(UC_ERR_PARSE, "KRS body:\n" + diag);
    auto s = std::make_unique<uc_system>();
    s->custom = true;
    s->device_source = ucuda::emit_custom_krs_source(krs_body, n_vars, n_values);
    s->n_vars = n_vars;
    s->n_values = n_values;
    *out = s.release();
    return UC_OK;
}

void uc_system_destroy(uc_system* s) { delete s; }
int uc_system_n_vars(const uc_system* s) { return s ? s->n_vars : 0; }
int uc_system_n_values(const uc_system* s) { return s ? s->n_values : 0; }

int uc_system_source(const uc_system* s, int dialect, char* buf, size_t cap, size_t* len) {
    if (!s) return UC_ERR_ARGUMENT;
    if (dialect == 0) return copy_out(program_source(s, false), buf, cap, len);
    if (s->custom) return UC_ERR_ARGUMENT;
    try {
        return copy_out(ucuda::emit_system_source(s->sys, ucuda::Dialect::C), buf, cap, len);
    } catch (...) {
        return UC_ERR_PARSE;
    }
}

int uc_system_op_counts(const uc_system* s, int* rhs_flops, int* rhs_calls, int* jac_nonzeros, int* jac_flops) {
    if (!s) return UC_ERR_ARGUMENT;
    if (rhs_flops) *rhs_flops = s->ops.rhs_flops;
    if (rhs_calls) *rhs_calls = s->ops.rhs_calls;
    if (jac_nonzeros) *jac_nonzeros = s->ops.jac_nonzeros;
    if (jac_flops) *jac_flops = s->ops.jac_flops;
    return UC_OK;
}

int uc_system_compile(uc_engine* e, uc_system* s, int strict, size_t* code_size) {
    if (!e || !s) return UC_ERR_ARGUMENT;
    precompile(e, s, strict != 0);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    Module* m = nullptr;
    const bool load = e->device >= 0;  // an engine with a device also loads the module
    int rc = load ? enter(e) : UC_OK;
    if (rc) return rc;
    rc = get_module(e, s, strict != 0, load, &m);
    if (rc) return rc;
    if (code_size) *code_size = m->code_size;
    return UC_OK;
}

int uc_codegen_scheme(uc_engine* e, const uc_system_desc* desc, int scheme, char* buf, size_t cap, size_t* len) {
    if (!desc) return UC_ERR_ARGUMENT;
    try {
        ucuda::System s = to_system(desc);
        return copy_out(ucuda::codegen_scheme(s, (ucuda::Scheme)scheme), buf, cap, len);
    } catch (const std::exception& ex) {
        if (e) e->fail(UC_ERR_PARSE, ex.what());
        return UC_ERR_PARSE;
    }
}

int uc_system_eval_rhs(const uc_system* s, const double* X, const double* a, double t, double* dX) {
    if (!s || !X || !a || !dX || s->custom) return UC_ERR_ARGUMENT;
    try {
        ucuda::eval_rhs(s->sys, X, a, t, dX);
    } catch (...) {
        return UC_ERR_PARSE;
    }
    return UC_OK;
}

long long uc_grid_cells(const uc_common* c) { return c ? cells_of(*c) : 0; }

int uc_axis_values(const uc_axis* ax, double* out) {
    if (!ax || !out || ax->n_pts <= 0) return UC_ERR_ARGUMENT;
    axis_values(*ax, out);
    return UC_OK;
}

// Validates a bifurcation request and fills the analysis-specific fields.
static int prep_bif(uc_engine* e, uc_system* s, const uc_bif_request* r, const uc_bif_output* o, long long* iters_out) {
    int rc = validate_common(e, s, r->c, true);
    if (rc) return rc;
    if (r->writable_var < -1 || r->writable_var >= s->n_vars) return e->fail(UC_ERR_ARGUMENT, "writable_var out of range");
    if (r->pre_scaller <= 0) return e->fail(UC_ERR_ARGUMENT, "pre_scaller must be > 0");
    if (r->max_peaks < 0) return e->fail(UC_ERR_ARGUMENT, "max_peaks must be >= 0");
    if (!o->flags || (r->max_peaks > 0 && (!o->peaks || !o->intervals)))
        return e->fail(UC_ERR_ARGUMENT, "bifurcation outputs missing");
    if (o->traj_len > 0 && !o->traj) return e->fail(UC_ERR_ARGUMENT, "traj_len > 0 but traj is NULL");
    // Recorded samples: ceil(t_max / h / pre_scaller) (run_bif1d, parametric_engine.cpp:1796),
    // transient steps: (int)(transient_time / h) (:1797).
    const double hw = step_of(r->c);
    if (!(hw > 0.0)) return e->fail(UC_ERR_ARGUMENT, "h must be > 0 (whole swept range)");
    const long long iters = (long long)std::ceil(r->c.t_max / hw / r->pre_scaller);
    if (iters <= 0) return e->fail(UC_ERR_ARGUMENT, "computed amountOfPointsInBlock <= 0");
    if (r->max_peaks > 2500) return e->fail(UC_ERR_ARGUMENT, "max_peaks above the reference cap 2500");
    if (r->record_precision != UC_PREC_F64 && r->record_precision != UC_PREC_F32C)
        return e->fail(UC_ERR_ARGUMENT, "unknown record_precision");
    *iters_out = iters;
    return UC_OK;
}

static void fill_bif(const uc_bif_request* r, const uc_bif_output* o, long long iters, UcJob& j) {
    j.record_f64 = r->c.precision == UC_PREC_F32C && r->record_precision == UC_PREC_F64;
    j.skip_steps = r->c.h > 0.0 ? (long long)(r->c.transient_time / r->c.h) : 0;
    j.n_iters = iters;
    j.pre_scaller = r->pre_scaller;
    j.writable_var = r->writable_var;
    j.max_peaks = r->max_peaks;
    j.traj_len = o->traj ? o->traj_len : 0;
    j.flag = o->flags;
    j.out0 = o->peaks;
    j.out1 = o->intervals;
    j.traj = o->traj;
    j.final_state = o->final_state;
}

int uc_run_bif(uc_engine* e, uc_system* s, const uc_bif_request* r, const uc_bif_output* o, void* stream) {
    if (!e || !s || !r || !o) return UC_ERR_ARGUMENT;
    precompile(e, s, r->c.strict != 0);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    long long iters = 0;
    int rc = prep_bif(e, s, r, o, &iters);
    if (rc) return rc;
    if ((rc = enter(e))) return rc;
    Module* m = nullptr;
    if ((rc = get_module(e, s, r->c.strict != 0, true, &m))) return rc;
    UcJob j;
    TableBuffer* tb = nullptr;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = stage_tables(e, s, r->c, j, st, &tb))) return rc;
    fill_bif(r, o, iters, j);
    return launch(e, m, r->c.precision == UC_PREC_F64 ? "uc_bif_f64" : "uc_bif_f32c", j, st, tb);
}

static int prep_lyap(uc_engine* e, uc_system* s, const uc_lyap_request* r, const uc_lyap_output* o) {
    int rc = validate_common(e, s, r->c, true);
    if (rc) return rc;
    if (!(r->NT > 0.0)) return e->fail(UC_ERR_ARGUMENT, "NT must be > 0");
    if (!(r->eps > 0.0)) return e->fail(UC_ERR_ARGUMENT, "eps must be > 0");
    if (r->method != UC_METHOD_VARIATIONAL && r->method != UC_METHOD_FINITE_DIFF) return e->fail(UC_ERR_ARGUMENT, "unknown method");
    if (s->custom && r->method != UC_METHOD_FINITE_DIFF)
        return e->fail(UC_ERR_ARGUMENT, "a custom step body has no Jacobian: use the finite-difference method");
    if (!o->values || !o->flags) return e->fail(UC_ERR_ARGUMENT, "outputs missing");
    // NT blocks: (int)(t_max / NT); steps per block (int)(NT / h); transient
    // (int)(transient_time / h) (run_lle_2d, parametric_engine.cpp:2731-2732, 2865).
    const long long blocks = (long long)(r->c.t_max / r->NT);
    if (blocks <= 0) return e->fail(UC_ERR_ARGUMENT, "computed t_max / NT <= 0");
    bool step_axis = false;
    for (int i = 0; i < r->c.n_axes; ++i) step_axis |= r->c.axis[i].kind == UC_AXIS_STEP;
    const long long nts = r->c.h > 0.0 ? (long long)(r->NT / r->c.h) : 0;
    if (!step_axis && nts <= 0) return e->fail(UC_ERR_ARGUMENT, "computed NT / h <= 0");
    return UC_OK;
}

static void fill_lyap(const uc_lyap_request* r, const uc_lyap_output* o, UcJob& j) {
    j.skip_steps = r->c.h > 0.0 ? (long long)(r->c.transient_time / r->c.h) : 0;
    j.n_iters = (long long)(r->c.t_max / r->NT);
    j.nt_steps = r->c.h > 0.0 ? (long long)(r->NT / r->c.h) : 0;
    j.nt = r->NT;
    j.eps = r->eps;
    j.method = r->method;
    j.probe_seed = r->probe_seed;
    j.flag = o->flags;
    j.out0 = o->values;
    j.final_state = o->final_state;
}

static const char* lyap_kernel(bool spectrum, int precision) {
    return spectrum ? (precision == UC_PREC_F64 ? "uc_ls_f64" : "uc_ls_f32c")
                    : (precision == UC_PREC_F64 ? "uc_lle_f64" : "uc_lle_f32c");
}

// Cells per lane of a LLE map on the lean loop (uc_lle_*_x1 / uc_lle_*_x2),
// or 0 for the general kernel. The lean loop needs fast mode, one tangent
// vector by the variational method, wave-uniform counts (no sweep over h)
// and blocks of 1..100 steps; it gives the general kernel's bits. Two cells
// per lane where that measured faster, with pw = pair-lane waves per SIMD:
// FP32C for 1.5 < pw <= 2 (-6% at 512^2; 4-9% slower at pw 2.75..4.9, even
// at 4), FP64 for pw >= 0.95 (-3..-7% from 362^2 to 1024^2). Below one wave
// per SIMD a lone wave cannot hide latency even with two cells (1.6x slower
// at 256^2).
static int lle_lanes(uc_engine* e, bool spectrum, const uc_lyap_request& r, bool step_axis, long long cells) {
    if (spectrum || r.c.strict || r.method != UC_METHOD_VARIATIONAL || step_axis) return 0;
    const long long nts = r.c.h > 0.0 ? (long long)(r.NT / r.c.h) : 0;
    if (nts < 1 || nts > 100) return 0;
    const double pw = (double)((cells + 1) / 2) / ((double)cu_count(e->device) * 4.0 * 64.0);
    const bool two = r.c.precision == UC_PREC_F64 ? pw >= 0.95 : (pw > 1.5 && pw <= 2.0);
    return two ? 2 : 1;
}

// Whether a spectrum map runs its N tangent vectors on N lanes per cell
// (uc_ls_*_xn): variational method, counts the same for every cell (no
// sweep over h), N >= 2, and a grid with fewer one-lane waves per SIMD than
// UC_SPREAD_BELOW, where one lane per cell leaves the SIMDs latency-bound.
// Measured (hyper-Rossler, 1.1e4 steps): up to 128^2 cells (0.25 waves) N
// lanes take 3.1 ms against 6.1 ms; at 181^2 (0.5) 3.9 / 5.6 ms (FP32C /
// FP64) against 6.2 / 6.0 ms; at 256^2 (1.0) 7.6 / 10.6 ms against 6.2 /
// 6.5 ms; Lorenz alike (3.1 against 4.9 ms on small grids).
#ifndef UC_SPREAD_BELOW
#define UC_SPREAD_BELOW 0.6
#endif
static bool ls_spread(uc_engine* e, const uc_lyap_request& r, int n_vars, bool step_axis, long long cells) {
    if (r.method != UC_METHOD_VARIATIONAL || step_axis || n_vars < 2) return false;
    const double w1 = (double)cells / ((double)cu_count(e->device) * 4.0 * 64.0);
    return w1 < UC_SPREAD_BELOW;
}

// Launch of the spread spectrum kernel: floor(64 / N) cells per wave.
static int launch_spread(uc_engine* e, Module* m, int precision, int n_vars, UcJob& j, hipStream_t st,
                         TableBuffer* tb) {
    const long long per = 64 / n_vars;
    const long long waves = (j.n_cells + per - 1) / per;
    return launch(e, m, precision == UC_PREC_F64 ? "uc_ls_f64_xn" : "uc_ls_f32c_xn", j, st, tb, waves * 64);
}

// The LLE kernel a lane rule picks, and its lane count.
static int launch_lle(uc_engine* e, Module* m, int lanes, int precision, UcJob& j, hipStream_t st, TableBuffer* tb) {
    const std::string name = std::string(lyap_kernel(false, precision)) + (lanes == 2 ? "_x2" : "_x1");
    return launch(e, m, name.c_str(), j, st, tb, lanes == 2 ? (j.n_cells + 1) / 2 : j.n_cells);
}

static int run_lyap(uc_engine* e, uc_system* s, const uc_lyap_request* r, const uc_lyap_output* o, void* stream, bool spectrum) {
    if (!e || !s || !r || !o) return UC_ERR_ARGUMENT;
    precompile(e, s, r->c.strict != 0);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    int rc = prep_lyap(e, s, r, o);
    if (rc) return rc;
    if ((rc = enter(e))) return rc;
    Module* m = nullptr;
    if ((rc = get_module(e, s, r->c.strict != 0, true, &m))) return rc;
    UcJob j;
    TableBuffer* tb = nullptr;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = stage_tables(e, s, r->c, j, st, &tb))) return rc;
    fill_lyap(r, o, j);
    bool step_axis = false;
    for (int i = 0; i < r->c.n_axes; ++i) step_axis |= r->c.axis[i].kind == UC_AXIS_STEP;
    // the lean loop (one or two cells per lane) for fast single-vector
    // variational maps with blocks of <= 100 steps
    if (const int lanes = lle_lanes(e, spectrum, *r, step_axis, j.n_cells))
        return launch_lle(e, m, lanes, r->c.precision, j, st, tb);
    // spectrum maps on small grids: one lane per tangent vector
    if (spectrum && ls_spread(e, *r, s->n_vars, step_axis, j.n_cells))
        return launch_spread(e, m, r->c.precision, s->n_vars, j, st, tb);
    return launch(e, m, lyap_kernel(spectrum, r->c.precision), j, st, tb);
}

int uc_run_lle(uc_engine* e, uc_system* s, const uc_lyap_request* r, const uc_lyap_output* o, void* stream) {
    return run_lyap(e, s, r, o, stream, false);
}

int uc_run_ls(uc_engine* e, uc_system* s, const uc_lyap_request* r, const uc_lyap_output* o, void* stream) {
    return run_lyap(e, s, r, o, stream, true);
}

// uc_run_lle with the launch tables in the GPU's static scratch array
// (scratch.hip) instead of a table buffer from the allocator: no hipMalloc,
// hipFree, hipHostMalloc or module load happens here. The module must be
// loaded already (uc_system_compile on this engine, or any earlier run of
// the system with the same strict flag); the pinned host image was made with
// the engine. Same kernels and launch shape as uc_run_lle, so the results
// are identical. Calls that overlap on one GPU share its scratch and race.
int uc_run_lle_noalloc(uc_engine* e, uc_system* s, const uc_lyap_request* r, const uc_lyap_output* o, void* stream) {
    if (!e || !s || !r || !o) return UC_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    int rc = prep_lyap(e, s, r, o);
    if (rc) return rc;
    const size_t need = table_doubles(r->c);
    if (need > (size_t)UC_SCRATCH_DOUBLES)
        return e->fail(UC_ERR_ARGUMENT, "uc_run_lle_noalloc: the axes need " + std::to_string(need) +
                                            " doubles of scratch, the maximum is " + std::to_string(UC_SCRATCH_DOUBLES) +
                                            " (nx + ny <= " + std::to_string(UC_SCRATCH_DOUBLES - kMaxValues - kMaxVars) + ")");
    if ((rc = enter(e))) return rc;
    if (!e->noalloc_error.empty()) return e->fail(UC_ERR_HIP, "uc_run_lle_noalloc: " + e->noalloc_error);
    double* dev = (double*)uc_scratch_ptr(e->device);
    if (!dev || !e->noalloc_host || !e->noalloc_done) return e->fail(UC_ERR_HIP, "uc_run_lle_noalloc: scratch not set up");
    Module* m = find_module(e, module_key(program_source(s, r->c.strict != 0), r->c.strict != 0));
    if (!m || !m->mod)
        return e->fail(UC_ERR_ARGUMENT, "uc_run_lle_noalloc: the system's module is not loaded on this engine "
                                        "(call uc_system_compile with the same strict flag first)");
    hipStream_t st = (hipStream_t)stream;
    // the previous call's copy must have read the host image before it is rewritten
    hipError_t he = hipEventSynchronize(e->noalloc_done);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "hipEventSynchronize(scratch staging)"));
    fill_tables(s->n_vars, r->c, e->noalloc_host);
    he = hipMemcpyAsync(dev, e->noalloc_host, need * sizeof(double), hipMemcpyHostToDevice, st);
    if (he == hipSuccess) he = hipEventRecord(e->noalloc_done, st);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "hipMemcpyAsync(scratch tables)"));
    UcJob j;
    wire_job(r->c, dev, j);
    fill_lyap(r, o, j);
    bool step_axis = false;
    for (int i = 0; i < r->c.n_axes; ++i) step_axis |= r->c.axis[i].kind == UC_AXIS_STEP;
    if (const int lanes = lle_lanes(e, false, *r, step_axis, j.n_cells))
        return launch_lle(e, m, lanes, r->c.precision, j, st, nullptr);
    return launch(e, m, lyap_kernel(false, r->c.precision), j, st, nullptr);
}

}  // extern "C"

namespace {

// Request checks shared by uc_run_basins and the server; returns the
// engine status and the number of labelled variables.
int prep_basins(uc_engine* e, uc_system* s, const uc_basins_request* r, uc_common& c, int& lbl) {
    if (s->custom) return e->fail(UC_ERR_ARGUMENT, "basins need a right-hand side (adaptive integrator), not a step body");
    c = r->c;
    if (c.t_max <= 0.0) c.t_max = 1.0;  // unused by basins
    int rc = validate_common(e, s, c, false);
    if (rc) return rc;
    if ((rc = whole_grid_only(e, c, "basins"))) return rc;
    lbl = s->n_vars < 2 ? s->n_vars : 2;
    if (!(r->rtol > 0.0) || !(r->atol > 0.0)) return e->fail(UC_ERR_ARGUMENT, "rtol and atol must be > 0");
    if (!(r->section_period > 0.0)) return e->fail(UC_ERR_ARGUMENT, "section_period must be > 0");
    if (r->n_sections < 1) return e->fail(UC_ERR_ARGUMENT, "n_sections must be >= 1");
    if (r->n_record < 2 || r->n_record > 16) return e->fail(UC_ERR_ARGUMENT, "n_record must be in [2, 16]");
    if (r->max_period < 1 || r->max_period > 8 || r->max_period >= r->n_record)
        return e->fail(UC_ERR_ARGUMENT, "max_period must be in [1, min(8, n_record - 1)]");
    if (r->label_vars != 0 && r->label_vars != lbl) return e->fail(UC_ERR_ARGUMENT, "label_vars must be min(n_vars, 2)");
    if (!(r->period_tol > 0.0) || !(r->quantum > 0.0)) return e->fail(UC_ERR_ARGUMENT, "period_tol and quantum must be > 0");
    if (!(r->h_min > 0.0) || !(r->h_max >= r->h_min)) return e->fail(UC_ERR_ARGUMENT, "need 0 < h_min <= h_max");
    return UC_OK;
}

void fill_basins(const uc_basins_request* r, int lbl, UcJob& j) {
    j.rtol = r->rtol; j.atol = r->atol; j.h_min = r->h_min; j.h_max = r->h_max;
    j.section_period = r->section_period; j.n_sections = r->n_sections; j.n_record = r->n_record;
    j.max_period = r->max_period; j.label_vars = lbl; j.period_tol = r->period_tol; j.quantum = r->quantum;
    j.max_attempts = r->max_attempts > 0 ? r->max_attempts : 100000000LL;
}

struct BasinsCatalogue {
    int n_attractors = 0;
    std::vector<int> period;      // per attractor, canonical order
    std::vector<double> point;    // [2] per attractor
};

// Attractor catalogue of a finished basins integration and the per-cell
// labels: the distinct quantised signatures go into a device hash set
// (growing it while it overflows), the host clusters keys of equal period
// whose quantised points are adjacent, orders the clusters by (period,
// point) with the no-cycle class last and numbers them 1..K, and a kernel
// writes every cell's label. `j` is the whole-grid descriptor (flag =
// periods, out0 = canonical points, label = labels). Device temporaries come
// from alloc / release; returns an error text, empty on success.
std::string basins_catalogue(Module* m, UcJob& j, double quantum, hipStream_t st,
                             const std::function<hipError_t(void**, size_t)>& alloc,
                             const std::function<void(void*)>& release, BasinsCatalogue& cat) {
    std::vector<void*> temps;
    auto cleanup = [&]() { for (void* p : temps) release(p); temps.clear(); };
    auto dev_alloc = [&](void** p, size_t bytes) -> bool {
        if (alloc(p, bytes) != hipSuccess) return false;
        temps.push_back(*p);
        return true;
    };
    const long long n = j.n_cells;
    hipFunction_t fkeys, flabels;
    hipError_t he = hipModuleGetFunction(&fkeys, m->mod, "uc_bas_keys");
    if (he == hipSuccess) he = hipModuleGetFunction(&flabels, m->mod, "uc_bas_labels");
    if (he != hipSuccess) return hip_msg(he, "basins label kernels");
    const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    int cap = 1 << 16;
    std::vector<unsigned long long> keys;
    unsigned long long* table = nullptr;
    unsigned long long* overflow = nullptr;
    if (!dev_alloc((void**)&overflow, sizeof(unsigned long long))) { cleanup(); return "basins: device memory"; }
    for (;;) {
        if (!dev_alloc((void**)&table, sizeof(unsigned long long) * (size_t)cap)) { cleanup(); return "basins: device memory"; }
        he = hipMemsetAsync(table, 0xFF, sizeof(unsigned long long) * (size_t)cap, st);
        if (he == hipSuccess) he = hipMemsetAsync(overflow, 0, sizeof(unsigned long long), st);
        void* kargs[] = {&j, &table, &cap, &overflow};
        if (he == hipSuccess) he = hipModuleLaunchKernel(fkeys, blocks, 1, 1, kBlock, 1, 1, 0, st, kargs, nullptr);
        unsigned long long ovf = 0;
        keys.assign((size_t)cap, 0ULL);
        if (he == hipSuccess) he = hipMemcpyAsync(keys.data(), table, sizeof(unsigned long long) * (size_t)cap, hipMemcpyDeviceToHost, st);
        if (he == hipSuccess) he = hipMemcpyAsync(&ovf, overflow, sizeof ovf, hipMemcpyDeviceToHost, st);
        if (he == hipSuccess) he = hipStreamSynchronize(st);
        if (he != hipSuccess) { cleanup(); return hip_msg(he, "basins signature keys"); }
        if (ovf == 0) break;
        if (cap >= (1 << 24)) { cleanup(); return "too many distinct attractor signatures (raise quantum)"; }
        cap <<= 2;
    }

    // ---- host: cluster adjacent keys, canonical order, labels 1..K ----
    struct KeyInfo { int p; long long qx, qv; };
    auto decode = [](unsigned long long k) {
        KeyInfo ki;
        ki.p = (int)(k >> 56);
        auto sx = [](unsigned long long v) { v &= (1ULL << 28) - 1ULL; return (long long)(v ^ (1ULL << 27)) - (1LL << 27); };
        ki.qx = sx(k >> 28);
        ki.qv = sx(k);
        return ki;
    };
    std::vector<int> slots;
    std::unordered_map<unsigned long long, int> index;
    for (int i = 0; i < cap; ++i)
        if (keys[(size_t)i] != ~0ULL) { index[keys[(size_t)i]] = (int)slots.size(); slots.push_back(i); }
    std::vector<int> parent(slots.size());
    for (size_t i = 0; i < parent.size(); ++i) parent[i] = (int)i;
    std::function<int(int)> find = [&](int x) { while (parent[(size_t)x] != x) x = parent[(size_t)x] = parent[(size_t)parent[(size_t)x]]; return x; };
    auto encode = [](int p, long long qx, long long qv) {
        const unsigned long long mask = (1ULL << 28) - 1ULL;
        return ((unsigned long long)(p & 0xFF) << 56) | (((unsigned long long)qx & mask) << 28) | ((unsigned long long)qv & mask);
    };
    for (size_t i = 0; i < slots.size(); ++i) {
        KeyInfo ki = decode(keys[(size_t)slots[i]]);
        if (ki.p == 0) continue;
        for (long long dx = -1; dx <= 1; ++dx)
            for (long long dv = -1; dv <= 1; ++dv) {
                auto it = index.find(encode(ki.p, ki.qx + dx, ki.qv + dv));
                if (it != index.end()) { int a = find((int)i), b = find(it->second); if (a != b) parent[(size_t)std::max(a, b)] = std::min(a, b); }
            }
    }
    // representative per cluster: the smallest (qx, qv) key
    std::map<int, KeyInfo> rep;
    for (size_t i = 0; i < slots.size(); ++i) {
        KeyInfo ki = decode(keys[(size_t)slots[i]]);
        int root = find((int)i);
        auto it = rep.find(root);
        if (it == rep.end() || std::make_pair(ki.qx, ki.qv) < std::make_pair(it->second.qx, it->second.qv)) rep[root] = ki;
    }
    std::vector<std::pair<std::tuple<int, long long, long long>, int>> order;
    for (auto& kv : rep) order.push_back({std::make_tuple(kv.second.p == 0 ? 1000 : kv.second.p, kv.second.qx, kv.second.qv), kv.first});
    std::sort(order.begin(), order.end());
    std::map<int, int> label_of_root;
    for (size_t k = 0; k < order.size(); ++k) label_of_root[order[k].second] = (int)k + 1;
    std::vector<int> slot_label((size_t)cap, 0);
    for (size_t i = 0; i < slots.size(); ++i) slot_label[(size_t)slots[i]] = label_of_root[find((int)i)];
    cat.n_attractors = (int)order.size();
    cat.period.clear();
    cat.point.clear();
    for (size_t k = 0; k < order.size(); ++k) {
        const KeyInfo& ki = rep[order[k].second];
        cat.period.push_back(ki.p);
        cat.point.push_back(ki.p ? (double)ki.qx * quantum : 0.0);
        cat.point.push_back(ki.p ? (double)ki.qv * quantum : 0.0);
    }
    int* dlabels = nullptr;
    if (!dev_alloc((void**)&dlabels, sizeof(int) * (size_t)cap)) { cleanup(); return "basins: device memory"; }
    he = hipMemcpyAsync(dlabels, slot_label.data(), sizeof(int) * (size_t)cap, hipMemcpyHostToDevice, st);
    void* largs[] = {&j, &table, &dlabels, &cap};
    if (he == hipSuccess) he = hipModuleLaunchKernel(flabels, blocks, 1, 1, kBlock, 1, 1, 0, st, largs, nullptr);
    if (he == hipSuccess) he = hipStreamSynchronize(st);
    cleanup();
    return he == hipSuccess ? std::string() : hip_msg(he, "basins labels");
}

void copy_catalogue(const BasinsCatalogue& cat, const uc_basins_output* o) {
    if (o->n_attractors) *o->n_attractors = cat.n_attractors;
    for (int k = 0; k < cat.n_attractors && k < o->max_attractors; ++k) {
        if (o->attractor_period) o->attractor_period[k] = cat.period[(size_t)k];
        if (o->attractor_point) {
            o->attractor_point[2 * k] = cat.point[2 * (size_t)k];
            o->attractor_point[2 * k + 1] = cat.point[2 * (size_t)k + 1];
        }
    }
}

}  // namespace

extern "C" {

int uc_run_basins(uc_engine* e, uc_system* s, const uc_basins_request* r, const uc_basins_output* o, void* stream) {
    if (!e || !s || !r || !o) return UC_ERR_ARGUMENT;
    precompile(e, s, r->c.strict != 0);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    uc_common c;
    int lbl = 0;
    int rc = prep_basins(e, s, r, c, lbl);
    if (rc) return rc;
    if (!o->labels || !o->periods) return e->fail(UC_ERR_ARGUMENT, "labels / periods outputs missing");
    if ((rc = enter(e))) return rc;

    Module* m = nullptr;
    if ((rc = get_module(e, s, c.strict != 0, true, &m))) return rc;
    UcJob j;
    TableBuffer* tb = nullptr;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = stage_tables(e, s, c, j, st, &tb))) return rc;
    const long long n = j.n_cells;
    fill_basins(r, lbl, j);
    j.flag = o->periods;
    j.counters = o->steps;
    j.label = o->labels;

    // canonical points: the caller's buffer, or a temporary
    double* points = o->cycle_point;
    if (!points) {
        hipError_t he = hipMalloc((void**)&points, sizeof(double) * (size_t)lbl * (size_t)n);
        if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "hipMalloc(basins)"));
    }
    auto free_points = [&]() { if (points != o->cycle_point) (void)hipFree(points); };
    j.out0 = points;
    if ((rc = launch(e, m, c.precision == UC_PREC_F64 ? "uc_bas_f64" : "uc_bas_f32c", j, st, tb))) { free_points(); return rc; }
    BasinsCatalogue cat;
    const std::string why = basins_catalogue(
        m, j, r->quantum, st, [](void** p, size_t b) { return hipMalloc(p, b); }, [](void* p) { (void)hipFree(p); }, cat);
    free_points();
    if (!why.empty()) return e->fail(why.rfind("too many", 0) == 0 ? UC_ERR_ARGUMENT : UC_ERR_HIP, why);
    copy_catalogue(cat, o);
    return UC_OK;
}

// Peak clustering launches shared by uc_run_bif_periods and
// uc_bif_dbscan_periods: the wave-per-cell kernels for cells with at most
// kDbNarrowMax peaks (and the 0 / 1 / -1 counts) and with at most
// kDbWaveMax, then the lane-per-cell kernel (HBM scratch labels / slots of
// max_n rows) for the others; a kernel is skipped when no cell can reach
// its range.
static const int kDbNarrowMax = 512;  // UC_DB_NARROW_MAX in the kernel source
static const int kDbWaveMax = 1024;   // UC_DB_WAVE_MAX in the kernel source

static int launch_dbscan(uc_engine* e, Module* m, const double* peaks, const int* counts, long long stride,
                         long long n_cells, int max_n, double eps, int* labels, int* slots, int* out, hipStream_t st) {
    hipFunction_t fw, fww, fl;
    hipError_t he = hipModuleGetFunction(&fw, m->mod, "uc_bif_dbscan_wave");
    if (he == hipSuccess) he = hipModuleGetFunction(&fww, m->mod, "uc_bif_dbscan_wave_wide");
    if (he == hipSuccess) he = hipModuleGetFunction(&fl, m->mod, "uc_bif_dbscan");
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "uc_bif_dbscan"));
    void* wargs[] = {(void*)&peaks, (void*)&counts, &stride, &n_cells, &max_n, &eps, &out};
    const long long waves_per_block = kBlock / 64;
    const unsigned wave_blocks = (unsigned)((n_cells + waves_per_block - 1) / waves_per_block);
    he = hipModuleLaunchKernel(fw, wave_blocks, 1, 1, kBlock, 1, 1, 0, st, wargs, nullptr);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "uc_bif_dbscan_wave launch"));
    if (max_n <= kDbNarrowMax) return UC_OK;
    he = hipModuleLaunchKernel(fww, wave_blocks, 1, 1, kBlock, 1, 1, 0, st, wargs, nullptr);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "uc_bif_dbscan_wave_wide launch"));
    if (max_n <= kDbWaveMax) return UC_OK;
    int min_n = kDbWaveMax;
    void* largs[] = {(void*)&peaks, (void*)&counts, &stride, &n_cells, &max_n, &eps, &min_n, &labels, &slots, &out};
    he = hipModuleLaunchKernel(fl, (unsigned)((n_cells + kBlock - 1) / kBlock), 1, 1, kBlock, 1, 1, 0, st, largs,
                               nullptr);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "uc_bif_dbscan launch"));
    return UC_OK;
}

// 2-D bifurcation periods (run_bif2d, U-CUDA/parametric_engine.cpp:5487-5911):
// the bifurcation sweep keeping every peak of the recorded window (the
// reference clusters all of them), then the per-cell DBSCAN kernel. Cells
// run in chunks sized so that the peaks and the clustering scratch fit in a
// quarter of the free HBM.
int uc_run_bif_periods(uc_engine* e, uc_system* s, const uc_bif_request* r, double eps, int* periods, void* stream) {
    if (!e || !s || !r) return UC_ERR_ARGUMENT;
    precompile(e, s, r->c.strict != 0);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    if (!periods) return e->fail(UC_ERR_ARGUMENT, "periods output missing");
    if (!(eps > 0.0)) return e->fail(UC_ERR_ARGUMENT, "eps_dbscan must be > 0");
    uc_bif_request rq = *r;
    rq.max_peaks = 0;
    int dflag = 0;
    uc_bif_output probe{&dflag, nullptr, nullptr, nullptr, 0, nullptr};
    long long iters = 0;
    int rc = prep_bif(e, s, &rq, &probe, &iters);
    if (rc) return rc;
    if ((rc = whole_grid_only(e, r->c, "uc_run_bif_periods"))) return rc;
    if ((rc = enter(e))) return rc;
    Module* m = nullptr;
    if ((rc = get_module(e, s, r->c.strict != 0, true, &m))) return rc;
    UcJob j;
    TableBuffer* tb = nullptr;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = stage_tables(e, s, r->c, j, st, &tb))) return rc;
    fill_bif(&rq, &probe, iters, j);
    // Raw peaks of `iters` samples are >= 2 apart: at most cap per cell. The
    // first pass keeps up to cap1 = kDbWaveMax + 1 (the wave kernel's limit
    // plus one); a chunk where some cell reports cap1 peaks may hold more,
    // and is swept again with the full capacity, in sub-chunks, for the lane
    // kernel (the sweep is deterministic, so its peaks are the same). This
    // keeps the common case at ~8 KB of peaks per cell instead of ~80 KB.
    const long long cap = iters / 2 + 2;
    if (cap > 0x7fffffffLL) return e->fail(UC_ERR_ARGUMENT, "recorded window too long");
    const long long cap1 = std::min<long long>(cap, kDbWaveMax + 1);
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    const long long total = j.n_cells;
    const long long chunk = (long long)std::min<size_t>(
        (size_t)total, std::max<size_t>(kBlock, free_b / 4 / ((size_t)cap1 * sizeof(double) + sizeof(int))));
    double *peaks = nullptr, *peaks_full = nullptr;
    int *flags = nullptr, *labels = nullptr, *slots = nullptr;
    long long full_chunk = 0;
    auto release = [&]() {
        (void)hipStreamSynchronize(st);
        for (void* p : {(void*)peaks, (void*)peaks_full, (void*)flags, (void*)labels, (void*)slots})
            if (p) (void)hipFree(p);
    };
    hipError_t he = hipMalloc((void**)&peaks, sizeof(double) * (size_t)cap1 * (size_t)chunk);
    if (he == hipSuccess) he = hipMalloc((void**)&flags, sizeof(int) * (size_t)chunk);
    if (he != hipSuccess) { release(); return e->fail(UC_ERR_HIP, hip_msg(he, "hipMalloc(bif periods)")); }
    const char* kname = r->c.precision == UC_PREC_F64 ? "uc_bif_f64" : "uc_bif_f32c";
    auto sweep = [&](long long off, long long n, double* pk, long long mp) {
        UcJob jc = j;
        jc.cell_offset = off;
        jc.n_cells = n;
        jc.stride = n;
        jc.flag = flags;
        jc.out0 = pk;
        jc.out1 = nullptr;
        jc.max_peaks = (int)mp;
        jc.traj = nullptr;
        jc.traj_len = 0;
        jc.final_state = nullptr;
        return launch(e, m, kname, jc, st, tb);
    };
    std::vector<int> hflags;
    for (long long off = 0; off < total; off += chunk) {
        const long long mcells = std::min(chunk, total - off);
        if ((rc = sweep(off, mcells, peaks, cap1))) { release(); return rc; }
        // a count above kDbWaveMax (only possible when cap1 > kDbWaveMax) needs
        // the lane kernel and its scratch: that chunk takes the full path
        bool overflow = false;
        if (cap1 > kDbWaveMax) {
            hflags.resize((size_t)mcells);
            he = hipMemcpyAsync(hflags.data(), flags, sizeof(int) * (size_t)mcells, hipMemcpyDeviceToHost, st);
            if (he == hipSuccess) he = hipStreamSynchronize(st);
            if (he != hipSuccess) { release(); return e->fail(UC_ERR_HIP, hip_msg(he, "bif periods counts")); }
            for (int f : hflags) overflow |= f > kDbWaveMax;
        }
        if (!overflow) {
            // every count <= kDbWaveMax here, so only the wave kernel runs
            const int max_n = (int)std::min<long long>(cap1, kDbWaveMax);
            if ((rc = launch_dbscan(e, m, peaks, flags, mcells, mcells, max_n, eps, nullptr, nullptr, periods + off,
                                    st))) {
                release();
                return rc;
            }
            continue;
        }
        if (!peaks_full) {
            const size_t per_cell = (size_t)cap * (sizeof(double) + 2 * sizeof(int));
            full_chunk = (long long)std::min<size_t>((size_t)chunk, std::max<size_t>(kBlock, free_b / 4 / per_cell));
            he = hipMalloc((void**)&peaks_full, sizeof(double) * (size_t)cap * (size_t)full_chunk);
            if (he == hipSuccess) he = hipMalloc((void**)&labels, sizeof(int) * (size_t)cap * (size_t)full_chunk);
            if (he == hipSuccess) he = hipMalloc((void**)&slots, sizeof(int) * (size_t)cap * (size_t)full_chunk);
            if (he != hipSuccess) { release(); return e->fail(UC_ERR_HIP, hip_msg(he, "hipMalloc(bif periods, full)")); }
        }
        for (long long sub = off; sub < off + mcells; sub += full_chunk) {
            const long long ns = std::min(full_chunk, off + mcells - sub);
            if ((rc = sweep(sub, ns, peaks_full, cap))) { release(); return rc; }
            if ((rc = launch_dbscan(e, m, peaks_full, flags, ns, ns, (int)cap, eps, labels, slots, periods + sub,
                                    st))) {
                release();
                return rc;
            }
        }
    }
    release();
    return UC_OK;
}

static const int kDftCells = 8;    // UC_DFT_C in the kernel source: cells per DFT workgroup
static const int kDftSamples = 4;  // samples per load block of uc_dft_cells (traj / window rows padded to it)

// Per-cell DFT tables for blocks of per-cell length (the h-sweep and the
// continuation DFT, dft1d_hsweep.template.cu / dft1d_cont.template.cu): block
// (int)(t_max / h / pre) capped at max_n, the window of that length (one per
// distinct length, cpu_build_window), rotation (cos, sin)(2 pi h pre f_k)
// and float-precision resync values (cpu_dft_block's formulas, C library).
struct DftCellTables {
    std::vector<int> Ns;
    std::vector<long long> win_off, res_off;
    std::vector<double> win, tab, res;
};

static void dft_cell_tables(const uc_dft_request* r, const std::vector<double>& h_cell, long long max_n,
                            DftCellTables& T) {
    const double kPi = 3.14159265358979323846;
    const int nf = r->n_freq;
    const size_t total = h_cell.size();
    T.Ns.assign(total, 0);
    T.win_off.assign(total, 0);
    T.res_off.assign(total, 0);
    T.win.clear();
    T.res.clear();
    T.tab.assign(2 * total * (size_t)nf, 0.0);
    std::map<long long, long long> win_of_len;
    std::vector<double> fk((size_t)nf);
    for (int k = 0; k < nf; ++k) {
        if (r->freq_log_scale) {
            const double l0 = std::log10(r->freq_lo), l1 = std::log10(r->freq_hi);
            fk[(size_t)k] = std::pow(10.0, l0 + (nf > 1 ? (l1 - l0) * (double)k / (double)(nf - 1) : 0.0));
        } else {
            const double f_step = nf > 1 ? (r->freq_hi - r->freq_lo) / (double)(nf - 1) : 0.0;
            fk[(size_t)k] = r->freq_lo + (double)k * f_step;
        }
    }
    for (size_t cell = 0; cell < total; ++cell) {
        const double h = h_cell[cell];
        long long n = h > 0.0 ? (long long)(r->c.t_max / h / (double)r->pre_scaller) : 0;
        if (n > max_n) n = max_n;
        T.Ns[cell] = (int)n;
        if (n <= 2) continue;  // unusable: the recording kernel flags it -1
        auto it = win_of_len.find(n);
        if (it == win_of_len.end()) {
            const long long off = (long long)T.win.size();
            const long long padded = (n + kDftSamples - 1) / kDftSamples * kDftSamples;
            T.win.resize((size_t)(off + padded), 0.0);
            if (r->window_type == 0) {
                std::fill(T.win.begin() + off, T.win.begin() + off + n, 1.0);
            } else {
                const double gamma = 2.0 * kPi / (double)(n - 1);
                for (long long q = 0; q < n; ++q)
                    T.win[(size_t)(off + q)] = r->window_type == 2 ? 0.53836 - 0.46164 * std::cos(gamma * (double)q)
                                                                   : 0.5 * (1.0 - std::cos(gamma * (double)q));
            }
            it = win_of_len.emplace(n, off).first;
        }
        T.win_off[cell] = it->second;
        const double dt = h * (double)r->pre_scaller;
        const double psi = 2.0 * kPi * dt;
        const int n_res = (int)((n - 1) / 1000);
        T.res_off[cell] = (long long)T.res.size();
        T.res.resize(T.res.size() + 2 * (size_t)nf * (size_t)n_res);
        double* rs = T.res.data() + T.res_off[cell];
        for (int k = 0; k < nf; ++k) {
            const double f = fk[(size_t)k];
            T.tab[2 * (cell * nf + k)] = std::cos(2.0 * kPi * dt * f);
            T.tab[2 * (cell * nf + k) + 1] = std::sin(2.0 * kPi * dt * f);
            for (int q = 0; q < n_res; ++q) {
                const double exact = psi * f * (double)(1000LL * (q + 1));
                rs[2 * ((size_t)k * n_res + q)] = (double)std::cos((float)exact);
                rs[2 * ((size_t)k * n_res + q) + 1] = (double)std::sin((float)exact);
            }
        }
    }
    if (T.win.empty()) T.win.assign(kDftSamples, 0.0);
    if (T.res.empty()) T.res.assign(2, 0.0);
}

// Device copies of the tables; `sums` launches uc_dft_cells over cells
// [off, off + n) of a traj buffer with plane stride `stride`.
struct DftCellTablesDev {
    double *win = nullptr, *tab = nullptr, *res = nullptr;
    int* N = nullptr;
    long long *wo = nullptr, *ro = nullptr;
    void free_all() {
        for (void* p : {(void*)win, (void*)tab, (void*)res, (void*)N, (void*)wo, (void*)ro})
            if (p) (void)hipFree(p);
    }
    hipError_t upload(const DftCellTables& T, hipStream_t st) {
        auto up = [&](void** d, const void* h, size_t bytes) {
            hipError_t he = hipMalloc(d, bytes);
            if (he == hipSuccess) he = hipMemcpyAsync(*d, h, bytes, hipMemcpyHostToDevice, st);
            return he;
        };
        hipError_t he = up((void**)&win, T.win.data(), sizeof(double) * T.win.size());
        if (he == hipSuccess) he = up((void**)&tab, T.tab.data(), sizeof(double) * T.tab.size());
        if (he == hipSuccess) he = up((void**)&res, T.res.data(), sizeof(double) * T.res.size());
        if (he == hipSuccess) he = up((void**)&N, T.Ns.data(), sizeof(int) * T.Ns.size());
        if (he == hipSuccess) he = up((void**)&wo, T.win_off.data(), sizeof(long long) * T.win_off.size());
        if (he == hipSuccess) he = up((void**)&ro, T.res_off.data(), sizeof(long long) * T.res_off.size());
        return he;
    }
    hipError_t sums(hipFunction_t f, double* traj, long long stride, long long off, long long n, int nf,
                    const uc_dft_output* o, hipStream_t st) {
        const int* Np = N + off;
        int* fl = o->flags + off;
        const long long* wop = wo + off;
        const double* tp = tab + 2 * off * nf;
        const long long* rop = ro + off;
        double* A = o->ak_cos + off * nf;
        double* B = o->bk_sin + off * nf;
        int nfa = nf;
        void* args[] = {&traj, &stride, &n, &Np, &fl, &win, &wop, &nfa, &tp, &res, &rop, &A, &B};
        const long long lanes = n * ((nf + 63) / 64) * 64;
        return hipModuleLaunchKernel(f, (unsigned)((lanes + kBlock - 1) / kBlock), 1, 1, kBlock, 1, 1, 0, st, args,
                                     nullptr);
    }
};

// DFT under an h sweep (run_dft1d_hsweep_gpu,
// U-CUDA/parametric_engine.cpp:4272-4520, and its kernel
// dft1d_hsweep.template.cu): every cell has its own step, block length
// (int)(t_max / h / pre) capped at ceil(t_max / min(h lo, h hi) / pre),
// window and sampling interval. The recording kernel runs with raw_flags 2,
// the sums with uc_dft_cells over the per-cell tables.
static int run_dft_hsweep(uc_engine* e, uc_system* s, const uc_dft_request* r, const uc_dft_output* o, int step_ax,
                          hipStream_t st) {
    const uc_common& c = r->c;
    const uc_axis& sa = c.axis[step_ax];
    if (!(sa.lo > 0.0 && sa.hi > 0.0)) return e->fail(UC_ERR_ARGUMENT, "h lo / hi must be > 0 for an h sweep");
    if (r->n_freq < 1) return e->fail(UC_ERR_ARGUMENT, "n_freq must be > 0");
    if (!(r->freq_hi > r->freq_lo)) return e->fail(UC_ERR_ARGUMENT, "freq_hi must be > freq_lo");
    const double max_d = std::ceil(c.t_max / std::min(sa.lo, sa.hi) / (double)r->pre_scaller);
    if (!(max_d >= 1.0)) return e->fail(UC_ERR_ARGUMENT, "block of t_max / h / pre_scaller samples is empty");
    if (max_d > 2.0e9) return e->fail(UC_ERR_ARGUMENT, "h lo too small: block longer than 2e9 samples");
    const long long max_n = (long long)max_d;
    const int nf = r->n_freq;
    const int nx = c.axis[0].n_pts, ny = c.n_axes > 1 ? c.axis[1].n_pts : 1;
    const long long total = (long long)nx * ny;
    std::vector<double> hv((size_t)sa.n_pts), h_cell((size_t)total);
    axis_values(sa, hv.data());
    for (long long cell = 0; cell < total; ++cell)
        h_cell[(size_t)cell] = hv[(size_t)(step_ax == 0 ? cell % nx : (cell / nx) % ny)];
    DftCellTables T;
    dft_cell_tables(r, h_cell, max_n, T);

    int rc = enter(e);
    if (rc) return rc;
    Module* m = nullptr;
    if ((rc = get_module(e, s, c.strict != 0, true, &m))) return rc;
    UcJob j;
    TableBuffer* tb = nullptr;
    if ((rc = stage_tables(e, s, c, j, st, &tb))) return rc;
    const long long rows = (max_n + kDftSamples - 1) / kDftSamples * kDftSamples;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    const long long chunk =
        (long long)std::min<size_t>((size_t)total, std::max<size_t>(kBlock, free_b / 4 / ((size_t)rows * sizeof(double))));
    double* traj = nullptr;
    DftCellTablesDev D;
    auto release = [&]() {
        (void)hipStreamSynchronize(st);
        if (traj) (void)hipFree(traj);
        D.free_all();
    };
    hipError_t he = hipMalloc((void**)&traj, sizeof(double) * (size_t)rows * (size_t)chunk);
    if (he == hipSuccess) he = D.upload(T, st);
    if (he != hipSuccess) { release(); return e->fail(UC_ERR_HIP, hip_msg(he, "DFT h-sweep buffers")); }
    hipFunction_t fdft;
    he = hipModuleGetFunction(&fdft, m->mod, "uc_dft_cells");
    if (he != hipSuccess) { release(); return e->fail(UC_ERR_HIP, hip_msg(he, "uc_dft_cells")); }
    const char* kname = c.precision == UC_PREC_F64 ? "uc_bif_f64" : "uc_bif_f32c";
    for (long long off = 0; off < total; off += chunk) {
        long long mcells = std::min(chunk, total - off);
        UcJob jc = j;
        jc.cell_offset = off;
        jc.n_cells = mcells;
        jc.stride = mcells;
        jc.n_iters = max_n;
        jc.pre_scaller = r->pre_scaller;
        jc.writable_var = r->writable_var;
        jc.max_peaks = 0;
        jc.raw_flags = 2;
        jc.flag = o->flags + off;
        jc.traj = traj;
        jc.traj_len = 0;
        if ((rc = launch(e, m, kname, jc, st, tb))) { release(); return rc; }
        he = D.sums(fdft, traj, mcells, off, mcells, nf, o, st);
        if (he != hipSuccess) { release(); return e->fail(UC_ERR_HIP, hip_msg(he, "uc_dft_cells launch")); }
    }
    release();
    return UC_OK;
}

// 1-D DFT sweep: the bifurcation kernel records each cell's block (loop
// flag kept raw), the DFT kernel turns it into n_freq coefficients. The
// window, the per-frequency rotation and the float-precision resync values
// are computed here with the formulas of run_dft1d_classical / DFT_custom,
// so the device work is multiply-add only. An axis over h takes the h-sweep
// path above.
int uc_run_dft(uc_engine* e, uc_system* s, const uc_dft_request* r, const uc_dft_output* o, void* stream) {
    if (!e || !s || !r || !o) return UC_ERR_ARGUMENT;
    precompile(e, s, r->c.strict != 0);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    int rc = validate_common(e, s, r->c, true);
    if (rc) return rc;
    if ((rc = whole_grid_only(e, r->c, "uc_run_dft"))) return rc;
    if (!o->ak_cos || !o->bk_sin || !o->flags) return e->fail(UC_ERR_ARGUMENT, "DFT outputs missing");
    if (r->writable_var < -1 || r->writable_var >= s->n_vars) return e->fail(UC_ERR_ARGUMENT, "writable_var out of range");
    if (r->pre_scaller <= 0) return e->fail(UC_ERR_ARGUMENT, "pre_scaller must be > 0");
    if (r->window_type < 0 || r->window_type > 2) return e->fail(UC_ERR_ARGUMENT, "window_type must be 0, 1 or 2");
    if (r->freq_log_scale && !(r->freq_lo > 0.0 && r->freq_hi > 0.0))
        return e->fail(UC_ERR_ARGUMENT, "log frequency axis needs freq_lo, freq_hi > 0");
    for (int i = 0; i < r->c.n_axes; ++i)
        if (r->c.axis[i].kind == UC_AXIS_STEP) return run_dft_hsweep(e, s, r, o, i, (hipStream_t)stream);
    if (r->n_freq < 2) return e->fail(UC_ERR_ARGUMENT, "n_freq must be >= 2");
    if (!(r->c.h > 0.0)) return e->fail(UC_ERR_ARGUMENT, "h must be > 0");
    const long long N = (long long)(r->c.t_max / r->c.h / (double)r->pre_scaller);
    if (N < 2) return e->fail(UC_ERR_ARGUMENT, "block of t_max / h / pre_scaller samples must hold >= 2");
    if (N > 0x7fffffffLL) return e->fail(UC_ERR_ARGUMENT, "block too long");
    if ((rc = enter(e))) return rc;
    Module* m = nullptr;
    if ((rc = get_module(e, s, r->c.strict != 0, true, &m))) return rc;
    UcJob j;
    TableBuffer* tb = nullptr;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = stage_tables(e, s, r->c, j, st, &tb))) return rc;

    // host tables (formulas of build_window and DFT_custom)
    const double kPi = 3.14159265358979323846;
    const double dt = r->c.h * (double)r->pre_scaller;
    const long long n_rows = (N + kDftSamples - 1) / kDftSamples * kDftSamples;
    std::vector<double> window((size_t)n_rows, 0.0), tab(2 * (size_t)r->n_freq);
    if (r->window_type == 0) std::fill(window.begin(), window.begin() + N, 1.0);
    else {
        const double gamma = 2.0 * kPi / (double)(N - 1);
        for (long long n = 0; n < N; ++n)
            window[(size_t)n] = r->window_type == 2 ? 0.53836 - 0.46164 * std::cos(gamma * (double)n)
                                                    : 0.5 * (1.0 - std::cos(gamma * (double)n));
    }
    const int n_res = (int)((N - 1) / 1000);
    std::vector<double> resync(2 * (size_t)r->n_freq * (size_t)std::max(1, n_res));
    const double f_step = (r->freq_hi - r->freq_lo) / (double)(r->n_freq - 1);
    const double psi = 2.0 * kPi * dt;
    for (int k = 0; k < r->n_freq; ++k) {
        double fk;
        if (r->freq_log_scale) {
            const double l0 = std::log10(r->freq_lo), l1 = std::log10(r->freq_hi);
            fk = std::pow(10.0, l0 + (double)k * ((l1 - l0) / (double)(r->n_freq - 1)));
        } else {
            fk = r->freq_lo + (double)k * f_step;
        }
        tab[2 * (size_t)k] = std::cos(2.0 * kPi * dt * fk);
        tab[2 * (size_t)k + 1] = std::sin(2.0 * kPi * dt * fk);
        for (int q = 0; q < n_res; ++q) {
            const double exact = psi * fk * ((double)(1000LL * (q + 1) - 1) + 1.0);
            resync[2 * ((size_t)k * n_res + q)] = (double)std::cos((float)exact);
            resync[2 * ((size_t)k * n_res + q) + 1] = (double)std::sin((float)exact);
        }
    }
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    const long long total = j.n_cells;
    const long long chunk = (long long)std::min<size_t>((size_t)total, std::max<size_t>(kBlock, free_b / 4 / ((size_t)N * sizeof(double))));
    double *traj = nullptr, *dwin = nullptr, *dtab = nullptr, *dres = nullptr;
    auto release = [&]() {
        (void)hipStreamSynchronize(st);
        for (void* p : {(void*)traj, (void*)dwin, (void*)dtab, (void*)dres})
            if (p) (void)hipFree(p);
    };
    hipError_t he = hipMalloc((void**)&traj, sizeof(double) * ((size_t)n_rows * (size_t)chunk + kDftCells));
    if (he == hipSuccess) he = hipMalloc((void**)&dwin, sizeof(double) * window.size());
    if (he == hipSuccess) he = hipMalloc((void**)&dtab, sizeof(double) * tab.size());
    if (he == hipSuccess) he = hipMalloc((void**)&dres, sizeof(double) * resync.size());
    if (he == hipSuccess) he = hipMemcpyAsync(dwin, window.data(), sizeof(double) * window.size(), hipMemcpyHostToDevice, st);
    if (he == hipSuccess) he = hipMemcpyAsync(dtab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, st);
    if (he == hipSuccess) he = hipMemcpyAsync(dres, resync.data(), sizeof(double) * resync.size(), hipMemcpyHostToDevice, st);
    if (he != hipSuccess) { release(); return e->fail(UC_ERR_HIP, hip_msg(he, "DFT buffers")); }
    hipFunction_t fdft;
    he = hipModuleGetFunction(&fdft, m->mod, "uc_dft");
    if (he != hipSuccess) { release(); return e->fail(UC_ERR_HIP, hip_msg(he, "uc_dft")); }
    const char* kname = r->c.precision == UC_PREC_F64 ? "uc_bif_f64" : "uc_bif_f32c";
    for (long long off = 0; off < total; off += chunk) {
        long long mcells = std::min(chunk, total - off);
        UcJob jc = j;
        jc.cell_offset = off;
        jc.n_cells = mcells;
        jc.stride = mcells;
        jc.skip_steps = (long long)(r->c.transient_time / r->c.h);
        jc.n_iters = N;
        jc.pre_scaller = r->pre_scaller;
        jc.writable_var = r->writable_var;
        jc.max_peaks = 0;
        jc.raw_flags = 1;
        jc.flag = o->flags + off;
        jc.traj = traj;
        jc.traj_len = (int)N;
        if ((rc = launch(e, m, kname, jc, st, tb))) { release(); return rc; }
        int Ni = (int)N, nf = r->n_freq, nr = n_res;
        long long stride = mcells;
        int* fl = o->flags + off;
        double* A = o->ak_cos + off * nf;
        double* B = o->bk_sin + off * nf;
        void* args[] = {&traj, &stride, &mcells, &Ni, &fl, &dwin, &nf, &dtab, &dres, &nr, &A, &B};
        const long long wgs = (mcells + kDftCells - 1) / kDftCells * ((nf + kBlock - 1) / kBlock);
        he = hipModuleLaunchKernel(fdft, (unsigned)wgs, 1, 1, kBlock, 1, 1, 0, st, args, nullptr);
        if (he != hipSuccess) { release(); return e->fail(UC_ERR_HIP, hip_msg(he, "uc_dft launch")); }
    }
    release();
    return UC_OK;
}

// ---- continuation sweeps ----
// Chain points of axis 0 in visiting order (cont_sweep_value,
// U-CUDA/parametric_engine.cpp:177-186, the formula the continuation kernels
// inline): t = j / (n - 1), lo + (hi - lo) t, from hi when reversed; log:
// 10^(l0 + (l1 - l0) t).
static void cont_values(const uc_axis& ax, int reverse, double* out) {
    const int n = ax.n_pts;
    const double denom = (double)(n > 1 ? n - 1 : 1);
    for (int j = 0; j < n; ++j) {
        const double t = (double)j / denom;
        if (ax.log_scale) {
            const double l0 = std::log10(ax.lo), l1 = std::log10(ax.hi);
            out[j] = std::pow(10.0, reverse ? (l1 - (l1 - l0) * t) : (l0 + (l1 - l0) * t));
        } else {
            out[j] = reverse ? (ax.hi - (ax.hi - ax.lo) * t) : (ax.lo + (ax.hi - ax.lo) * t);
        }
    }
}

// Shared set-up: checks, module, tables, the chain table of axis 0 (owned
// by the caller: freed after the stream is synchronised), and the smallest
// step any chain point can take.
struct ContRun {
    Module* m = nullptr;
    UcJob j;
    TableBuffer* tb = nullptr;
    double* ax0 = nullptr;
    double min_h = 0.0;
};

static int prep_cont(uc_engine* e, uc_system* s, const uc_common& c, int reverse, hipStream_t st, ContRun& R) {
    int rc = validate_common(e, s, c, true);
    if (rc) return rc;
    if ((rc = whole_grid_only(e, c, "a continuation sweep"))) return rc;
    const uc_axis& a0 = c.axis[0];
    if (a0.kind != UC_AXIS_PARAM && a0.kind != UC_AXIS_STEP)
        return e->fail(UC_ERR_ARGUMENT, "continuation sweeps a parameter or h (axis 0), not an initial state");
    if (c.precision != UC_PREC_F64) return e->fail(UC_ERR_ARGUMENT, "continuation sweeps run in FP64");
    if (c.ic_seed != 0ULL && c.n_axes < 2) return e->fail(UC_ERR_ARGUMENT, "random initial states need a chain axis");
    auto hmin = [&](const uc_axis& ax) { return std::min(ax.lo, ax.hi); };
    R.min_h = a0.kind == UC_AXIS_STEP ? hmin(a0) : (c.n_axes > 1 && c.axis[1].kind == UC_AXIS_STEP ? hmin(c.axis[1]) : c.h);
    if (!(R.min_h > 0.0)) return e->fail(UC_ERR_ARGUMENT, "h must be > 0 (every point of an h sweep)");
    if ((rc = enter(e))) return rc;
    if ((rc = get_module(e, s, c.strict != 0, true, &R.m))) return rc;
    if ((rc = stage_tables(e, s, c, R.j, st, &R.tb))) return rc;
    std::vector<double> v((size_t)a0.n_pts);
    cont_values(a0, reverse, v.data());
    hipError_t he = hipMalloc((void**)&R.ax0, sizeof(double) * v.size());
    if (he == hipSuccess) he = hipMemcpyAsync(R.ax0, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, st);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "continuation table"));
    R.j.axis_vals[0] = R.ax0;
    return UC_OK;
}

static int launch_chains(uc_engine* e, ContRun& R, const char* name, hipStream_t st) {
    hipFunction_t f;
    hipError_t he = hipModuleGetFunction(&f, R.m->mod, name);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, name));
    void* args[] = {&R.j};
    he = hipModuleLaunchKernel(f, (unsigned)((R.j.ny + kBlock - 1) / kBlock), 1, 1, kBlock, 1, 1, 0, st, args, nullptr);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, name));
    he = hipEventRecord(R.tb->done, st);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "hipEventRecord"));
    R.tb->in_flight = true;
    return UC_OK;
}

static void cont_done(ContRun& R, hipStream_t st) {
    (void)hipStreamSynchronize(st);
    if (R.ax0) (void)hipFree(R.ax0);
    R.ax0 = nullptr;
}

int uc_run_bif_cont(uc_engine* e, uc_system* s, const uc_bif_request* r, int reverse, const uc_bif_output* o,
                    void* stream) {
    if (!e || !s || !r || !o) return UC_ERR_ARGUMENT;
    precompile(e, s, r->c.strict != 0);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    if (!o->flags) return e->fail(UC_ERR_ARGUMENT, "flags output missing");
    if (r->max_peaks < 0) return e->fail(UC_ERR_ARGUMENT, "max_peaks must be >= 0");
    if (r->max_peaks > 0 && (!o->peaks || !o->intervals)) return e->fail(UC_ERR_ARGUMENT, "peaks / intervals missing");
    if (r->writable_var < -1 || r->writable_var >= s->n_vars) return e->fail(UC_ERR_ARGUMENT, "writable_var out of range");
    if (r->pre_scaller <= 0) return e->fail(UC_ERR_ARGUMENT, "pre_scaller must be > 0");
    hipStream_t st = (hipStream_t)stream;
    ContRun R;
    int rc = prep_cont(e, s, r->c, reverse, st, R);
    if (!rc) {
        const double max_d = std::ceil(r->c.t_max / R.min_h / (double)r->pre_scaller);
        if (!(max_d >= 1.0) || max_d > 2.0e9) rc = e->fail(UC_ERR_ARGUMENT, "block of t_max / h / pre_scaller samples out of range");
        R.j.n_iters = (long long)max_d;
        R.j.pre_scaller = r->pre_scaller;
        R.j.writable_var = r->writable_var;
        R.j.max_peaks = r->max_peaks;
        R.j.flag = o->flags;
        R.j.out0 = r->max_peaks > 0 ? o->peaks : nullptr;
        R.j.out1 = r->max_peaks > 0 ? o->intervals : nullptr;
        R.j.raw_flags = 0;
        if (!rc) rc = launch_chains(e, R, "uc_cont_bif", st);
    }
    cont_done(R, st);
    return rc;
}

static int run_lyap_cont(uc_engine* e, uc_system* s, const uc_lyap_request* r, int reverse, const uc_lyap_output* o,
                         void* stream, bool spectrum) {
    if (!e || !s || !r || !o) return UC_ERR_ARGUMENT;
    precompile(e, s, r->c.strict != 0);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    if (!o->values || !o->flags) return e->fail(UC_ERR_ARGUMENT, "outputs missing");
    if (!(r->NT > 0.0)) return e->fail(UC_ERR_ARGUMENT, "NT must be > 0");
    if (!(r->eps > 0.0)) return e->fail(UC_ERR_ARGUMENT, "eps must be > 0");
    hipStream_t st = (hipStream_t)stream;
    ContRun R;
    int rc = prep_cont(e, s, r->c, reverse, st, R);
    if (!rc) {
        R.j.nt = r->NT;
        R.j.eps = r->eps;
        R.j.probe_seed = r->probe_seed;
        R.j.flag = o->flags;
        R.j.out0 = o->values;
        rc = launch_chains(e, R, spectrum ? "uc_cont_ls" : "uc_cont_lle", st);
    }
    cont_done(R, st);
    return rc;
}

int uc_run_lle_cont(uc_engine* e, uc_system* s, const uc_lyap_request* r, int reverse, const uc_lyap_output* o,
                    void* stream) {
    return run_lyap_cont(e, s, r, reverse, o, stream, false);
}

int uc_run_ls_cont(uc_engine* e, uc_system* s, const uc_lyap_request* r, int reverse, const uc_lyap_output* o,
                   void* stream) {
    return run_lyap_cont(e, s, r, reverse, o, stream, true);
}

int uc_run_dft_cont(uc_engine* e, uc_system* s, const uc_dft_request* r, int reverse, const uc_dft_output* o,
                    void* stream) {
    if (!e || !s || !r || !o) return UC_ERR_ARGUMENT;
    precompile(e, s, r->c.strict != 0);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    if (!o->ak_cos || !o->bk_sin || !o->flags) return e->fail(UC_ERR_ARGUMENT, "DFT outputs missing");
    if (r->writable_var < -1 || r->writable_var >= s->n_vars) return e->fail(UC_ERR_ARGUMENT, "writable_var out of range");
    if (r->pre_scaller <= 0) return e->fail(UC_ERR_ARGUMENT, "pre_scaller must be > 0");
    if (r->n_freq < 1) return e->fail(UC_ERR_ARGUMENT, "n_freq must be > 0");
    if (r->window_type < 0 || r->window_type > 2) return e->fail(UC_ERR_ARGUMENT, "window_type must be 0, 1 or 2");
    if (!(r->freq_hi > r->freq_lo)) return e->fail(UC_ERR_ARGUMENT, "freq_hi must be > freq_lo");
    if (r->freq_log_scale && !(r->freq_lo > 0.0 && r->freq_hi > 0.0))
        return e->fail(UC_ERR_ARGUMENT, "log frequency axis needs freq_lo, freq_hi > 0");
    hipStream_t st = (hipStream_t)stream;
    ContRun R;
    int rc = prep_cont(e, s, r->c, reverse, st, R);
    if (rc) { cont_done(R, st); return rc; }
    const uc_common& c = r->c;
    const double max_d = std::ceil(c.t_max / R.min_h / (double)r->pre_scaller);
    if (!(max_d >= 1.0) || max_d > 2.0e9) {
        cont_done(R, st);
        return e->fail(UC_ERR_ARGUMENT, "block of t_max / h / pre_scaller samples out of range");
    }
    const long long max_n = (long long)max_d;
    const int nx = c.axis[0].n_pts, ny = c.n_axes > 1 ? c.axis[1].n_pts : 1;
    const long long total = (long long)nx * ny;
    // every cell's step: the chain point (h axis 0), the chain (h axis 1) or h
    std::vector<double> v0((size_t)nx), v1((size_t)ny, c.h), h_cell((size_t)total);
    cont_values(c.axis[0], reverse, v0.data());
    if (c.n_axes > 1 && c.axis[1].kind == UC_AXIS_STEP) axis_values(c.axis[1], v1.data());
    for (long long cell = 0; cell < total; ++cell)
        h_cell[(size_t)cell] = c.axis[0].kind == UC_AXIS_STEP ? v0[(size_t)(cell % nx)] : v1[(size_t)(cell / nx)];
    DftCellTables T;
    dft_cell_tables(r, h_cell, max_n, T);
    const long long rows = (max_n + kDftSamples - 1) / kDftSamples * kDftSamples;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    if ((double)rows * (double)total * sizeof(double) > (double)free_b / 2) {
        cont_done(R, st);
        return e->fail(UC_ERR_ARGUMENT, "continuation DFT blocks exceed half of the free HBM");
    }
    double* traj = nullptr;
    DftCellTablesDev D;
    hipError_t he = hipMalloc((void**)&traj, sizeof(double) * ((size_t)rows * (size_t)total + kDftCells));
    if (he == hipSuccess) he = D.upload(T, st);
    if (he == hipSuccess) {
        R.j.n_iters = max_n;
        R.j.pre_scaller = r->pre_scaller;
        R.j.writable_var = r->writable_var;
        R.j.max_peaks = 0;
        R.j.flag = o->flags;
        R.j.traj = traj;
        R.j.raw_flags = 1;
        rc = launch_chains(e, R, "uc_cont_bif", st);
        hipFunction_t fdft;
        if (!rc) {
            he = hipModuleGetFunction(&fdft, R.m->mod, "uc_dft_cells");
            if (he == hipSuccess) he = D.sums(fdft, traj, total, 0, total, r->n_freq, o, st);
        }
    }
    if (!rc && he != hipSuccess) rc = e->fail(UC_ERR_HIP, hip_msg(he, "continuation DFT"));
    cont_done(R, st);
    if (traj) (void)hipFree(traj);
    D.free_all();
    return rc;
}

int uc_peak_finder(uc_engine* e, uc_system* s, const double* data, long long n_samples, long long n_series,
                   double time_step, int max_peaks, int* counts, double* peaks, double* intervals, void* stream) {
    if (!e || !s) return UC_ERR_ARGUMENT;
    precompile(e, s, true);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    if (!data || !counts || n_samples < 0 || n_series <= 0 || max_peaks < 0 || max_peaks > 2500)
        return e->fail(UC_ERR_ARGUMENT, "peak finder: bad arguments");
    if (max_peaks > 0 && (!peaks || !intervals)) return e->fail(UC_ERR_ARGUMENT, "peak finder: outputs missing");
    int rc = enter(e);
    if (rc) return rc;
    Module* m = nullptr;
    if ((rc = get_module(e, s, true, true, &m))) return rc;
    hipFunction_t f;
    hipError_t he = hipModuleGetFunction(&f, m->mod, "uc_peaks_series");
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "uc_peaks_series"));
    const unsigned blocks = (unsigned)((n_series + kBlock - 1) / kBlock);
    void* args[] = {(void*)&data, &n_samples, &n_series, &time_step, &max_peaks, &counts, &peaks, &intervals};
    he = hipModuleLaunchKernel(f, blocks, 1, 1, kBlock, 1, 1, 0, (hipStream_t)stream, args, nullptr);
    if (he != hipSuccess) return e->fail(UC_ERR_HIP, hip_msg(he, "hipModuleLaunchKernel(peaks)"));
    return UC_OK;
}

// ---- fast synchronisation ----
static int fs_validate(const uc_fastsync_request* r, std::string& why) {
    if (!r) { why = "request missing"; return UC_ERR_ARGUMENT; }
    if (!(r->h > 0.0)) { why = "h must be > 0"; return UC_ERR_ARGUMENT; }
    if (r->iter_of_synchr <= 0) { why = "iter_of_synchr must be > 0"; return UC_ERR_ARGUMENT; }
    if (r->pre_scaller <= 0) { why = "pre_scaller must be > 0"; return UC_ERR_ARGUMENT; }
    if (r->type_of_synch < 0 || r->type_of_synch > 1) { why = "type_of_synch must be 0 or 1"; return UC_ERR_ARGUMENT; }
    if (r->error_estim < 0 || r->error_estim > 3) { why = "error_estim must be 0..3"; return UC_ERR_ARGUMENT; }
    if (r->transient_time < 0.0) { why = "transient_time must be >= 0"; return UC_ERR_ARGUMENT; }
    if (r->mode == 0) {
        if (!(r->t_max > 0.0)) { why = "t_max must be > 0"; return UC_ERR_ARGUMENT; }
        if (!(r->window > 0.0)) { why = "window must be > 0"; return UC_ERR_ARGUMENT; }
        if ((int)(r->t_max / r->h) / r->pre_scaller <= 0) { why = "t_max / h / pre_scaller gives no points"; return UC_ERR_ARGUMENT; }
    } else if (r->mode == 1) {
        if (r->n_pts <= 0) { why = "n_pts must be > 0"; return UC_ERR_ARGUMENT; }
        if ((int)(r->window / r->h / (double)r->pre_scaller) <= 0) { why = "window / h / pre_scaller gives no samples"; return UC_ERR_ARGUMENT; }
    } else {
        why = "mode must be 0 (on attractor) or 1 (on grid)";
        return UC_ERR_ARGUMENT;
    }
    return UC_OK;
}

long long uc_fastsync_points(const uc_fastsync_request* r) {
    std::string why;
    if (fs_validate(r, why) != UC_OK) return -1;
    if (r->mode == 0) return (long long)((int)(r->t_max / r->h) / r->pre_scaller);
    return (long long)r->n_pts * r->n_pts;
}

// On attractor: one-lane master fill, then one lane per window. On grid: a
// one-lane transient of the fixed initial state, then chunks of cells (each
// unidirectional cell stores its master window, n_iters x n_vars doubles).
int uc_run_fastsync(uc_engine* e, uc_system* s, const uc_fastsync_request* r, const uc_fastsync_output* o,
                    void* stream) {
    if (!e || !s) return UC_ERR_ARGUMENT;
    if (r) precompile(e, s, r->strict != 0);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    std::string why;
    int rc = fs_validate(r, why);
    if (rc) return e->fail(rc, why);
    if (!o) return e->fail(UC_ERR_ARGUMENT, "output missing");
    const int N = s->n_vars;
    if (!r->k_forward || !r->k_backward || !r->ic_master || !r->ic_slave)
        return e->fail(UC_ERR_ARGUMENT, "k_forward / k_backward / ic_master / ic_slave missing");
    if (!r->values && r->n_values > 0) return e->fail(UC_ERR_ARGUMENT, "values is NULL");
    if (r->n_values > kMaxValues) return e->fail(UC_ERR_ARGUMENT, "too many values");
    if (r->n_values < s->n_values) return e->fail(UC_ERR_ARGUMENT, "values shorter than the system's a[] (a[0] + params)");
    if (!o->sync_error) return e->fail(UC_ERR_ARGUMENT, "sync_error output missing");
    if (s->device_source.find("#define UC_USES_TIME 1") != std::string::npos)
        return e->fail(UC_ERR_ARGUMENT, "fast synchronisation needs an autonomous system (no time symbol)");
    if (r->mode == 1) {
        if (r->axis_x_var < 0 || r->axis_x_var >= N || r->axis_y_var < 0 || r->axis_y_var >= N)
            return e->fail(UC_ERR_ARGUMENT, "axis_x_var / axis_y_var out of range");
        if (r->axis_x_var == r->axis_y_var) return e->fail(UC_ERR_ARGUMENT, "axis_x_var == axis_y_var");
    }
    if ((rc = enter(e))) return rc;
    Module* m = nullptr;
    if ((rc = get_module(e, s, r->strict != 0, true, &m))) return rc;
    hipStream_t st = (hipStream_t)stream;

    // small input tables: a[] (kMaxValues slots), ic_master, ic_slave,
    // k_forward, k_backward, grid axes
    const int np = r->mode == 1 ? r->n_pts : 1;
    std::vector<double> host((size_t)kMaxValues + 4 * (size_t)N + 2 * (size_t)np, 0.0);
    for (int i = 0; i < r->n_values; ++i) host[(size_t)i] = r->values[i];
    double* hp = host.data() + kMaxValues;
    for (int i = 0; i < N; ++i) {
        hp[i] = r->ic_master[i];
        hp[N + i] = r->ic_slave[i];
        hp[2 * N + i] = r->k_forward[i];
        hp[3 * N + i] = r->k_backward[i];
    }
    if (r->mode == 1) {
        uc_axis ax{};
        ax.n_pts = np;
        ax.lo = r->axis_x_lo;
        ax.hi = r->axis_x_hi;
        axis_values(ax, hp + 4 * N);
        ax.lo = r->axis_y_lo;
        ax.hi = r->axis_y_hi;
        axis_values(ax, hp + 4 * N + np);
    }
    double *tabs = nullptr, *buf = nullptr;
    auto release = [&]() {
        (void)hipStreamSynchronize(st);
        if (tabs) (void)hipFree(tabs);
        if (buf) (void)hipFree(buf);
    };
    hipError_t he = hipMalloc((void**)&tabs, sizeof(double) * host.size());
    if (he == hipSuccess) he = hipMemcpyAsync(tabs, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice, st);
    if (he != hipSuccess) { release(); return e->fail(UC_ERR_HIP, hip_msg(he, "fast synchro tables")); }
    UcFsJob J;
    std::memset(&J, 0, sizeof J);
    J.params = tabs;
    J.ic_master = tabs + kMaxValues;
    J.ic_slave = tabs + kMaxValues + N;
    J.k_forward = tabs + kMaxValues + 2 * N;
    J.k_backward = tabs + kMaxValues + 3 * N;
    J.h = r->h;
    J.iter_of_synchr = r->iter_of_synchr;
    J.pre_scaller = r->pre_scaller;
    J.type_of_synch = r->type_of_synch;
    J.error_estim = r->error_estim;
    J.fs_error_trs = r->fs_error_trs;
    auto get = [&](const char* name, hipFunction_t* f) {
        hipError_t x = hipModuleGetFunction(f, m->mod, name);
        if (x != hipSuccess) { release(); return e->fail(UC_ERR_HIP, hip_msg(x, name)); }
        return (int)UC_OK;
    };
    auto go = [&](hipFunction_t f, long long lanes, void** args, const char* what) {
        const long long blocks = (lanes + kBlock - 1) / kBlock;
        hipError_t x = hipModuleLaunchKernel(f, (unsigned)blocks, 1, 1, lanes == 1 ? 1 : kBlock, 1, 1, 0, st, args, nullptr);
        if (x != hipSuccess) { release(); return e->fail(UC_ERR_HIP, hip_msg(x, what)); }
        return (int)UC_OK;
    };
    const int skip = (int)(r->transient_time / r->h);
    if (r->mode == 0) {
        const int nt = (int)(r->window / r->h), ct = (int)(r->t_max / r->h);
        const long long n_out = ct / r->pre_scaller;
        const long long len = (long long)ct + nt;
        he = hipMalloc((void**)&buf, sizeof(double) * (size_t)std::max<long long>(len, 1) * (size_t)N);
        if (he != hipSuccess) { release(); return e->fail(UC_ERR_HIP, hip_msg(he, "hipMalloc(master trajectory)")); }
        hipFunction_t ff, fa;
        if ((rc = get("uc_fs_fill", &ff)) || (rc = get("uc_fs_attr", &fa))) return rc;
        int skip_a = skip, len_a = (int)len;
        void* fargs[] = {&J, &skip_a, &len_a, &buf};
        if ((rc = go(ff, 1, fargs, "uc_fs_fill"))) return rc;
        J.n_cells = n_out;
        J.n_iters = nt;
        J.master = buf;
        J.out = o->sync_error;
        J.traj_out = o->traj;
        void* aargs[] = {&J};
        if ((rc = go(fa, n_out, aargs, "uc_fs_attr"))) return rc;
        release();
        return UC_OK;
    }
    // on grid
    const int nb = (int)(r->window / r->h / (double)r->pre_scaller);
    const long long total = (long long)np * np;
    double* fixed = const_cast<double*>(r->grid_swap_master_slave ? J.ic_master : J.ic_slave);
    if (r->transient_time > 0.0 && skip > 0) {
        hipFunction_t ft;
        if ((rc = get("uc_fs_transient", &ft))) return rc;
        int skip_a = skip;
        void* targs[] = {&J, &skip_a, &fixed};
        if ((rc = go(ft, 1, targs, "uc_fs_transient"))) return rc;
    }
    long long chunk = total;
    if (r->type_of_synch == 0) {
        size_t free_b = 0, total_b = 0;
        (void)hipMemGetInfo(&free_b, &total_b);
        const size_t per_cell = (size_t)nb * (size_t)N * sizeof(double);
        chunk = (long long)std::min<size_t>((size_t)total, std::max<size_t>(kBlock, free_b / 4 / per_cell));
        he = hipMalloc((void**)&buf, per_cell * (size_t)chunk);
        if (he != hipSuccess) { release(); return e->fail(UC_ERR_HIP, hip_msg(he, "hipMalloc(master windows)")); }
    }
    hipFunction_t fg;
    if ((rc = get("uc_fs_grid", &fg))) return rc;
    J.n_pts = np;
    J.axis_var[0] = r->axis_x_var;
    J.axis_var[1] = r->axis_y_var;
    J.axis_vals[0] = tabs + kMaxValues + 4 * N;
    J.axis_vals[1] = tabs + kMaxValues + 4 * N + np;
    J.swap = r->grid_swap_master_slave ? 1 : 0;
    J.n_iters = nb;
    J.data = buf;
    for (long long off = 0; off < total; off += chunk) {
        UcFsJob Jc = J;
        Jc.cell_offset = off;
        Jc.n_cells = std::min(chunk, total - off);
        Jc.stride = Jc.n_cells;
        Jc.out = o->sync_error + off;
        void* gargs[] = {&Jc};
        if ((rc = go(fg, Jc.n_cells, gargs, "uc_fs_grid"))) return rc;
    }
    release();
    return UC_OK;
}

// Periods from stored peaks (dbscanCUDA, U-CUDA/cudaLibrary.cu:2321-2345).
int uc_bif_dbscan_periods(uc_engine* e, uc_system* s, const double* peaks, const int* counts, long long n_cells,
                          int max_peaks, double eps, int* periods, void* stream) {
    if (!e || !s) return UC_ERR_ARGUMENT;
    precompile(e, s, true);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    if (!counts || !periods || n_cells <= 0 || max_peaks < 0 || (max_peaks > 0 && !peaks))
        return e->fail(UC_ERR_ARGUMENT, "dbscan periods: bad arguments");
    if (!(eps > 0.0)) return e->fail(UC_ERR_ARGUMENT, "eps_dbscan must be > 0");
    int rc = enter(e);
    if (rc) return rc;
    Module* m = nullptr;
    if ((rc = get_module(e, s, true, true, &m))) return rc;
    hipStream_t st = (hipStream_t)stream;
    int *labels = nullptr, *slots = nullptr;
    hipError_t he = hipSuccess;
    if (max_peaks > kDbWaveMax) {
        he = hipMalloc((void**)&labels, sizeof(int) * (size_t)max_peaks * (size_t)n_cells);
        if (he == hipSuccess) he = hipMalloc((void**)&slots, sizeof(int) * (size_t)max_peaks * (size_t)n_cells);
    }
    if (he == hipSuccess)
        rc = launch_dbscan(e, m, peaks, counts, n_cells, n_cells, max_peaks, eps, labels, slots, periods, st);
    else
        rc = e->fail(UC_ERR_HIP, hip_msg(he, "hipMalloc(dbscan scratch)"));
    if (labels || slots) {
        (void)hipStreamSynchronize(st);
        if (labels) (void)hipFree(labels);
        if (slots) (void)hipFree(slots);
    }
    return rc;
}

}  // extern "C"

// =====================================================================
// Resumable slices. A job whose cells all step the same counts (no sweep
// over h) can run in slices: each launch of a sliced persistent kernel
// ("*_s") runs a range of every cell's work units and keeps unfinished
// cells in a slab of `words` doubles per cell. Units: transient chunks of
// UC_UNIT_STEPS steps, then one per renormalisation block (LLE / LS), per
// UC_UNIT_STEPS recorded samples (bifurcation) or per section (basins).
// Results equal a whole run's bit for bit.
namespace {
enum { kKindBif = 0, kKindLle = 1, kKindLs = 2, kKindBasins = 3 };
struct SliceInfo {
    long long units = 0;    // 0: cannot be sliced
    double unit_cost = 0.0; // traj-steps per unit, on average
    int words = 0;
};
long long unit_chunks(long long steps) { return steps <= 0 ? 1 : (steps + UC_UNIT_STEPS - 1) / UC_UNIT_STEPS; }
SliceInfo slice_info(int kind, const UcJob& j, int n_vars, double step_cost) {
    SliceInfo si;
    if (j.axis_kind[0] == UC_AXIS_STEP || j.axis_kind[1] == UC_AXIS_STEP || j.raw_flags == 2) return si;
    if (kind == kKindBif) {
        si.units = unit_chunks(j.skip_steps) + unit_chunks(j.n_iters);
        si.words = UC_SLAB_BIF_WORDS(n_vars);
    } else if (kind == kKindBasins) {
        si.units = j.n_sections;
        si.words = UC_SLAB_BAS_WORDS(n_vars);
    } else {
        si.units = unit_chunks(j.skip_steps) + j.n_iters;
        si.words = UC_SLAB_LYAP_WORDS(n_vars, kind == kKindLs ? n_vars : 1);
    }
    si.unit_cost = step_cost / (double)std::max(1LL, si.units);
    return si;
}
// units per launch for a target of `slice_steps` traj-steps per cell
long long slice_units(const SliceInfo& si, double slice_steps) {
    return std::max(1LL, (long long)(slice_steps / std::max(1e-9, si.unit_cost)));
}
std::string sliced_kernel(const std::string& q) { return q.substr(0, q.size() - 2) + "_s"; }  // "*_q" -> "*_s"
// adaptive basins: scheduling weight of one section in traj-steps (the
// measured FP32C / FP64 basin rates against the LLE map's)
double basins_section_cost(int precision) { return precision == UC_PREC_F64 ? 1024.0 : 256.0; }
}  // namespace

// =====================================================================
// Persistent heterogeneous batch: every job of the same (module, kernel)
// shares one persistent launch over a job table. A plan fixes the launch
// shape (block, resident waves per SIMD), the trajectory-steps per launch
// (0: whole runs; else the sliceable jobs run in stages of that many steps
// per cell, stage k + 1 after stage k) and the cells per lane (2: LLE maps
// that can run two cells per lane do). The reset kernel and every stage's
// launches are captured into one hipGraph and launched once.
struct uc_batch {
    uc_engine* e = nullptr;
    struct Job {
        Module* m;
        std::string kernel;  // persistent kernel for a whole run ("*_q")
        UcJob job;
        double cost;         // steps per cell, for longest-first ordering
        int kind;
        SliceInfo si;
        bool pair;           // can run the lean loop ("*_x1_q", two cells per lane "*_x2_q")
        double* slab;        // allocated by the first plan that slices the job
        int spread_n;        // N when a spectrum map can run one lane per tangent vector ("*_xn_q"), else 0
    };
    std::vector<Job> jobs;
    // basins jobs: the whole-grid descriptor for the catalogue and labels
    // built after each graph launch, and where the catalogue goes
    struct Basins {
        Module* m;
        UcJob job;
        double quantum;
        uc_basins_output out;
    };
    std::vector<Basins> basins;
    std::vector<void*> owned;  // device allocations (tables, job arrays, counters, slabs)
    struct Group {
        Module* m;
        std::string kernel;
        int cells_per_chunk_lane;  // 1, or 2 for the pair kernels
        int spread_n;              // N for the N-lanes-per-cell kernels (floor(64 / N) cells per wave), else 0
        std::vector<UcJob> table;
        UcJob* d_jobs = nullptr;
        long long* d_start = nullptr;
        long long total_chunks = 0;
    };
    std::vector<std::vector<Group>> stages;
    unsigned long long* d_counters = nullptr;
    int n_counters = 0;
    uc_batch_plan plan{256, 4, 0.0, 1, 0, 0.0};
    uc_batch_plan built{};  // plan the tables were built for (block 0: none)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipStream_t> aux;
    std::vector<hipEvent_t> events;

    ~uc_batch() {
        for (auto& j : jobs) --j.m->pins;
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto ev : events) (void)hipEventDestroy(ev);
        for (auto st : aux) (void)hipStreamDestroy(st);
        for (void* p : owned) (void)hipFree(p);
    }
    int fail(int code, const std::string& m) { return e->fail(code, m); }
    void drop_graph() {
        if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    }
    void invalidate() { built.block = 0; drop_graph(); }
};

namespace {

int batch_stage(uc_batch* b, const uc_system* s, const uc_common& c, UcJob& j) {
    const size_t need = table_doubles(c);
    std::vector<double> host(need);
    fill_tables(s->n_vars, c, host.data());
    double* dev = nullptr;
    hipError_t he = hipMalloc((void**)&dev, need * sizeof(double));
    if (he != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, "hipMalloc(batch tables)"));
    b->owned.push_back(dev);
    he = hipMemcpy(dev, host.data(), need * sizeof(double), hipMemcpyHostToDevice);
    if (he != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, "hipMemcpy(batch tables)"));
    wire_job(c, dev, j);
    return UC_OK;
}

int cu_count(int device) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || n <= 0) n = 256;
    return n;
}

void batch_add(uc_batch* b, Module* m, const std::string& kernel, const UcJob& j, double cost, int kind, int n_vars,
               bool pair, int spread_n = 0) {
    uc_batch::Job jb{m, kernel, j, cost, kind, slice_info(kind, j, n_vars, cost), pair, nullptr, spread_n};
    b->jobs.push_back(jb);
    ++m->pins;
    b->invalidate();
}

bool same_plan(const uc_batch_plan& a, const uc_batch_plan& p) {
    return a.block == p.block && a.slice_steps == p.slice_steps && a.cells_per_lane == p.cells_per_lane &&
           a.vector_lanes == p.vector_lanes;
}

// Builds the per-stage, per-group job tables for plan `p`.
int batch_build(uc_batch* b, const uc_batch_plan& p) {
    if (b->built.block && same_plan(b->built, p)) return UC_OK;
    b->invalidate();
    std::vector<const uc_batch::Job*> order;
    for (auto& jb : b->jobs) order.push_back(&jb);
    std::stable_sort(order.begin(), order.end(),
                     [](const uc_batch::Job* x, const uc_batch::Job* y) { return x->cost > y->cost; });
    b->stages.clear();
    for (const uc_batch::Job* jp : order) {
        uc_batch::Job& jb = const_cast<uc_batch::Job&>(*jp);
        std::string kernel = jb.kernel;
        int cpl = 1;
        long long n_stages = 1, ups = 0;
        int spread = 0;
        if (p.vector_lanes && jb.spread_n) {
            kernel = kernel.substr(0, kernel.size() - 2) + "_xn_q";  // one lane per tangent vector
            spread = jb.spread_n;
        } else if (p.cells_per_lane == 2 && jb.pair) {
            kernel = kernel.substr(0, kernel.size() - 2) + "_x2_q";
            cpl = 2;
        } else if (p.slice_steps > 0.0 && jb.si.units > 0 && (ups = slice_units(jb.si, p.slice_steps)) < jb.si.units) {
            kernel = sliced_kernel(kernel);
            n_stages = (jb.si.units + ups - 1) / ups;
            if (!jb.slab) {
                const size_t bytes = sizeof(double) * (size_t)jb.si.words * (size_t)jb.job.n_cells;
                hipError_t he = hipMalloc((void**)&jb.slab, bytes);
                if (he != hipSuccess) { jb.slab = nullptr; return b->fail(UC_ERR_HIP, hip_msg(he, "hipMalloc(batch slab)")); }
                b->owned.push_back(jb.slab);
            }
        } else if (jb.pair) {
            kernel = kernel.substr(0, kernel.size() - 2) + "_x1_q";  // the lean loop, one cell per lane
        }
        if ((long long)b->stages.size() < n_stages) b->stages.resize((size_t)n_stages);
        for (long long k = 0; k < n_stages; ++k) {
            UcJob j = jb.job;
            if (n_stages > 1) {
                j.slab = jb.slab;
                j.unit_begin = k * ups;
                j.unit_end = std::min(jb.si.units, (k + 1) * ups);
            }
            auto& st = b->stages[(size_t)k];
            size_t g = 0;
            while (g < st.size() && !(st[g].m == jb.m && st[g].kernel == kernel)) ++g;
            if (g == st.size()) st.push_back(uc_batch::Group{jb.m, kernel, cpl, spread, {}, nullptr, nullptr, 0});
            st[g].table.push_back(j);
        }
    }
    int n_launch = 0;
    for (auto& st : b->stages)
        for (auto& g : st) {
            std::vector<long long> start(g.table.size() + 1, 0);
            const long long chunk = g.spread_n ? (long long)(p.block / 64) * (64 / g.spread_n)
                                               : (long long)p.block * g.cells_per_chunk_lane;
            for (size_t i = 0; i < g.table.size(); ++i) start[i + 1] = start[i] + (g.table[i].n_cells + chunk - 1) / chunk;
            g.total_chunks = start.back();
            hipError_t he = hipMalloc((void**)&g.d_jobs, sizeof(UcJob) * g.table.size());
            if (he == hipSuccess) { b->owned.push_back(g.d_jobs); he = hipMalloc((void**)&g.d_start, sizeof(long long) * start.size()); }
            if (he == hipSuccess) { b->owned.push_back(g.d_start); he = hipMemcpy(g.d_jobs, g.table.data(), sizeof(UcJob) * g.table.size(), hipMemcpyHostToDevice); }
            if (he == hipSuccess) he = hipMemcpy(g.d_start, start.data(), sizeof(long long) * start.size(), hipMemcpyHostToDevice);
            if (he != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, "batch job tables"));
            ++n_launch;
        }
    if (n_launch > b->n_counters) {
        hipError_t he = hipMalloc((void**)&b->d_counters, sizeof(unsigned long long) * (size_t)n_launch);
        if (he != hipSuccess) { b->d_counters = nullptr; b->n_counters = 0; return b->fail(UC_ERR_HIP, hip_msg(he, "batch counters")); }
        b->owned.push_back(b->d_counters);
        b->n_counters = n_launch;
    }
    b->built = p;
    return UC_OK;
}

// Enqueues the reset kernel, then each stage: one persistent launch per
// group, the groups concurrently on auxiliary streams joined back into `st`
// before the next stage.
int batch_enqueue(uc_batch* b, hipStream_t st, const uc_batch_plan& p) {
    const int cus = cu_count(b->e->device);
    const int waves_per_wg = p.block / 64;
    long long wgs = (long long)cus * 4 * p.waves_per_simd / waves_per_wg;
    if (wgs < 1) wgs = 1;
    size_t max_groups = 0, n_launch = 0;
    for (auto& s : b->stages) { max_groups = std::max(max_groups, s.size()); n_launch += s.size(); }
    hipFunction_t freset;
    hipError_t he = hipModuleGetFunction(&freset, b->jobs[0].m->mod, "uc_batch_reset");
    int nc = (int)n_launch;
    void* rargs[] = {&b->d_counters, &nc};
    if (he == hipSuccess) he = hipModuleLaunchKernel(freset, (unsigned)((nc + 63) / 64), 1, 1, 64, 1, 1, 0, st, rargs, nullptr);
    if (he != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, "batch reset"));
    while (b->aux.size() + 1 < max_groups) {
        hipStream_t s2;
        if ((he = hipStreamCreateWithFlags(&s2, hipStreamNonBlocking)) != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, "aux stream"));
        b->aux.push_back(s2);
    }
    while (b->events.size() < 2 * n_launch) {
        hipEvent_t ev;
        if ((he = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, "event"));
        b->events.push_back(ev);
    }
    size_t ctr = 0, evi = 0;
    for (auto& stage : b->stages) {
        const int ng = (int)stage.size();
        hipEvent_t fork = b->events[evi++];
        if (ng > 1 && (he = hipEventRecord(fork, st)) != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, "fork"));
        for (int g = 0; g < ng; ++g) {
            auto& G = stage[(size_t)g];
            hipStream_t gs = g == 0 ? st : b->aux[(size_t)g - 1];
            if (g > 0 && (he = hipStreamWaitEvent(gs, fork, 0)) != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, "fork wait"));
            hipFunction_t f;
            if ((he = hipModuleGetFunction(&f, G.m->mod, G.kernel.c_str())) != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, G.kernel.c_str()));
            int nj = (int)G.table.size();
            unsigned long long* c = b->d_counters + ctr++;
            void* args[] = {&G.d_jobs, &G.d_start, &nj, &c};
            const long long grid = std::min<long long>(wgs, std::max<long long>(1, G.total_chunks));
            if ((he = hipModuleLaunchKernel(f, (unsigned)grid, 1, 1, (unsigned)p.block, 1, 1, 0, gs, args, nullptr)) != hipSuccess)
                return b->fail(UC_ERR_HIP, hip_msg(he, "persistent launch"));
            if (g > 0) {
                hipEvent_t join = b->events[evi++];
                if ((he = hipEventRecord(join, gs)) != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, "join"));
                if ((he = hipStreamWaitEvent(st, join, 0)) != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, "join wait"));
            }
        }
    }
    return UC_OK;
}

int check_plan(uc_batch* b, const uc_batch_plan& p) {
    if (p.block != 64 && p.block != 128 && p.block != 256) return b->fail(UC_ERR_ARGUMENT, "block must be 64, 128 or 256");
    if (p.waves_per_simd < 1 || p.waves_per_simd > 8) return b->fail(UC_ERR_ARGUMENT, "waves_per_simd must be in [1, 8]");
    if (!(p.slice_steps >= 0.0)) return b->fail(UC_ERR_ARGUMENT, "slice_steps must be >= 0");
    if (p.cells_per_lane != 1 && p.cells_per_lane != 2) return b->fail(UC_ERR_ARGUMENT, "cells_per_lane must be 1 or 2");
    if (p.vector_lanes != 0 && p.vector_lanes != 1) return b->fail(UC_ERR_ARGUMENT, "vector_lanes must be 0 or 1");
    return UC_OK;
}

}  // namespace

extern "C" {

int uc_batch_create(uc_engine* e, uc_batch** out) {
    if (!e || !out) return UC_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    int rc = enter(e);
    if (rc) return rc;
    auto b = new uc_batch;
    b->e = e;
    e->refs.fetch_add(1, std::memory_order_relaxed);
    *out = b;
    return UC_OK;
}

void uc_batch_destroy(uc_batch* b) {
    if (!b) return;
    uc_engine* e = b->e;
    {
        std::lock_guard<std::recursive_mutex> lk(e->mu);
        (void)hipSetDevice(e->device);
        delete b;
    }
    engine_release(e);
}

int uc_batch_add_bif(uc_batch* b, uc_system* s, const uc_bif_request* r, const uc_bif_output* o, int* job_id) {
    if (!b || !s || !r || !o) return UC_ERR_ARGUMENT;
    uc_engine* e = b->e;
    precompile(e, s, r->c.strict != 0);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    long long iters = 0;
    int rc = prep_bif(e, s, r, o, &iters);
    if (rc) return rc;
    if ((rc = whole_grid_only(e, r->c, "a batch job"))) return rc;
    if ((rc = enter(e))) return rc;
    Module* m = nullptr;
    if ((rc = get_module(e, s, r->c.strict != 0, true, &m))) return rc;
    UcJob j;
    if ((rc = batch_stage(b, s, r->c, j))) return rc;
    fill_bif(r, o, iters, j);
    const double cost = (double)(j.skip_steps + iters * r->pre_scaller);
    batch_add(b, m, r->c.precision == UC_PREC_F64 ? "uc_bif_f64_q" : "uc_bif_f32c_q", j, cost, kKindBif, s->n_vars,
              false);
    if (job_id) *job_id = (int)b->jobs.size() - 1;
    return UC_OK;
}

int uc_batch_add_lyap(uc_batch* b, uc_system* s, const uc_lyap_request* r, int spectrum, const uc_lyap_output* o,
                      int* job_id) {
    if (!b || !s || !r || !o) return UC_ERR_ARGUMENT;
    uc_engine* e = b->e;
    precompile(e, s, r->c.strict != 0);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    int rc = prep_lyap(e, s, r, o);
    if (rc) return rc;
    if ((rc = whole_grid_only(e, r->c, "a batch job"))) return rc;
    if ((rc = enter(e))) return rc;
    Module* m = nullptr;
    if ((rc = get_module(e, s, r->c.strict != 0, true, &m))) return rc;
    UcJob j;
    if ((rc = batch_stage(b, s, r->c, j))) return rc;
    fill_lyap(r, o, j);
    const double cost = (double)(j.skip_steps + j.n_iters * (j.nt_steps + 1)) * (spectrum ? 1.0 + s->n_vars : 2.0);
    // the lean loop (uc_lle_lanes, one or two cells per lane): the fast
    // single-vector variational map with wave-uniform counts and blocks of
    // 1..100 steps
    bool step_axis = false;
    for (int i = 0; i < r->c.n_axes; ++i) step_axis |= r->c.axis[i].kind == UC_AXIS_STEP;
    const bool pair = !spectrum && !r->c.strict && r->method == UC_METHOD_VARIATIONAL && !step_axis &&
                      j.nt_steps >= 1 && j.nt_steps <= 100;
    // one lane per tangent vector: spectrum maps by the variational method
    // with wave-uniform counts (uc_ls_spread)
    const int spread_n = spectrum && r->method == UC_METHOD_VARIATIONAL && !step_axis && s->n_vars >= 2 ? s->n_vars : 0;
    batch_add(b, m, std::string(lyap_kernel(spectrum != 0, r->c.precision)) + "_q", j, cost,
              spectrum ? kKindLs : kKindLle, s->n_vars, pair, spread_n);
    if (job_id) *job_id = (int)b->jobs.size() - 1;
    return UC_OK;
}

int uc_batch_add_basins(uc_batch* b, uc_system* s, const uc_basins_request* r, const uc_basins_output* o,
                        int* job_id) {
    if (!b || !s || !r || !o) return UC_ERR_ARGUMENT;
    uc_engine* e = b->e;
    precompile(e, s, r->c.strict != 0);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    uc_common c;
    int lbl = 0;
    int rc = prep_basins(e, s, r, c, lbl);
    if (rc) return rc;
    if (!o->labels || !o->periods) return e->fail(UC_ERR_ARGUMENT, "labels / periods outputs missing");
    if ((rc = enter(e))) return rc;
    Module* m = nullptr;
    if ((rc = get_module(e, s, c.strict != 0, true, &m))) return rc;
    UcJob j;
    if ((rc = batch_stage(b, s, c, j))) return rc;
    fill_basins(r, lbl, j);
    j.flag = o->periods;
    j.counters = o->steps;
    j.label = o->labels;
    double* points = o->cycle_point;
    if (!points) {
        hipError_t he = hipMalloc((void**)&points, sizeof(double) * (size_t)lbl * (size_t)j.n_cells);
        if (he != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, "hipMalloc(batch basins points)"));
        b->owned.push_back(points);
    }
    j.out0 = points;
    batch_add(b, m, c.precision == UC_PREC_F64 ? "uc_bas_f64_q" : "uc_bas_f32c_q", j,
              basins_section_cost(c.precision) * r->n_sections, kKindBasins, s->n_vars, false);
    b->basins.push_back({m, j, r->quantum, *o});
    if (job_id) *job_id = (int)b->jobs.size() - 1;
    return UC_OK;
}

int uc_batch_set_launch(uc_batch* b, int block, int waves_per_simd) {
    if (!b) return UC_ERR_ARGUMENT;
    uc_batch_plan p = b->plan;
    p.block = block;
    p.waves_per_simd = waves_per_simd;
    return uc_batch_set_plan(b, &p);
}

int uc_batch_set_plan(uc_batch* b, const uc_batch_plan* p) {
    if (!b || !p) return UC_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> lk(b->e->mu);
    int rc = check_plan(b, *p);
    if (rc) return rc;
    if (p->block != b->plan.block || p->waves_per_simd != b->plan.waves_per_simd ||
        p->slice_steps != b->plan.slice_steps || p->cells_per_lane != b->plan.cells_per_lane ||
        p->vector_lanes != b->plan.vector_lanes)
        b->drop_graph();
    b->plan = *p;
    b->plan.ms = 0.0;
    return UC_OK;
}

int uc_batch_get_plan(const uc_batch* b, uc_batch_plan* p) {
    if (!b || !p) return UC_ERR_ARGUMENT;
    *p = b->plan;
    return UC_OK;
}

// Times the batch (eagerly, without the graph) for each candidate plan and
// keeps the fastest: block {64, 256} x waves per SIMD {2, 4, 8} with whole
// runs, then at the best shape the steps per launch {0, 16384, 4096} and,
// when a job can run two cells per lane, cells per lane {1, 2}.
int uc_batch_tune_plan(uc_batch* b, void* stream, uc_batch_plan* best_out) {
    if (!b) return UC_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> lk(b->e->mu);
    if (b->jobs.empty()) return b->fail(UC_ERR_ARGUMENT, "empty batch");
    int rc = enter(b->e);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t t0, t1;
    if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) return b->fail(UC_ERR_HIP, "event create");
    uc_batch_plan best = b->plan;
    best.ms = 1e30;
    auto time_plan = [&](const uc_batch_plan& p) -> int {
        int r2 = batch_build(b, p);
        if (r2) return r2;
        (void)hipEventRecord(t0, st);
        if ((r2 = batch_enqueue(b, st, p))) return r2;
        (void)hipEventRecord(t1, st);
        if (hipEventSynchronize(t1) != hipSuccess) return b->fail(UC_ERR_HIP, "tune sync");
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, t0, t1);
        if (ms < best.ms) { best = p; best.ms = ms; }
        return UC_OK;
    };
    for (int blk : {64, 256})
        for (int w : {2, 4, 8})
            if (!rc) rc = time_plan(uc_batch_plan{blk, w, 0.0, 1, 0, 0.0});
    bool any_pair = false, any_slice = false, any_spread = false;
    for (auto& jb : b->jobs) { any_pair |= jb.pair; any_slice |= jb.si.units > 0; any_spread |= jb.spread_n > 0; }
    const uc_batch_plan shape = best;
    if (any_slice)
        for (double sl : {16384.0, 4096.0})
            if (!rc) rc = time_plan(uc_batch_plan{shape.block, shape.waves_per_simd, sl, 1, 0, 0.0});
    if (any_pair && !rc) rc = time_plan(uc_batch_plan{shape.block, shape.waves_per_simd, 0.0, 2, 0, 0.0});
    if (any_spread && !rc) {
        rc = time_plan(uc_batch_plan{shape.block, shape.waves_per_simd, 0.0, 1, 1, 0.0});
        if (any_pair && !rc) rc = time_plan(uc_batch_plan{shape.block, shape.waves_per_simd, 0.0, 2, 1, 0.0});
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc) return rc;
    b->drop_graph();
    b->plan = best;
    if (best_out) *best_out = best;
    return UC_OK;
}

int uc_batch_tune(uc_batch* b, void* stream, int* best_block, int* best_wps, double* best_ms) {
    uc_batch_plan p{};
    const int rc = uc_batch_tune_plan(b, stream, &p);
    if (rc) return rc;
    if (best_block) *best_block = p.block;
    if (best_wps) *best_wps = p.waves_per_simd;
    if (best_ms) *best_ms = p.ms;
    return UC_OK;
}

int uc_batch_launch(uc_batch* b, void* stream) {
    if (!b) return UC_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> lk(b->e->mu);
    if (b->jobs.empty()) return b->fail(UC_ERR_ARGUMENT, "empty batch");
    int rc = enter(b->e);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (!st) return b->fail(UC_ERR_ARGUMENT, "batch launch needs a non-default stream (graph capture)");
    const uc_batch_plan p = b->plan;
    if ((rc = batch_build(b, p))) return rc;
    if (!b->exec) {
        hipError_t he = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (he != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, "hipStreamBeginCapture"));
        rc = batch_enqueue(b, st, p);
        hipGraph_t g = nullptr;
        he = hipStreamEndCapture(st, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (he != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, "hipStreamEndCapture"));
        b->graph = g;
        he = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (he != hipSuccess) { b->exec = nullptr; return b->fail(UC_ERR_HIP, hip_msg(he, "hipGraphInstantiate")); }
    }
    hipError_t he = hipGraphLaunch(b->exec, st);
    if (he != hipSuccess) return b->fail(UC_ERR_HIP, hip_msg(he, "hipGraphLaunch"));
    // basins jobs: attractor catalogue and labels once the integration is
    // done (host clustering in between, so outside the graph; synchronises)
    for (auto& bj : b->basins) {
        BasinsCatalogue cat;
        const std::string why = basins_catalogue(
            bj.m, bj.job, bj.quantum, st, [](void** p2, size_t n) { return hipMalloc(p2, n); },
            [](void* p2) { (void)hipFree(p2); }, cat);
        if (!why.empty()) return b->fail(why.rfind("too many", 0) == 0 ? UC_ERR_ARGUMENT : UC_ERR_HIP, why);
        copy_catalogue(cat, &bj.out);
    }
    return UC_OK;
}

int uc_batch_info(const uc_batch* b, int* n_jobs, int* n_groups, int* block, int* waves_per_simd) {
    if (!b) return UC_ERR_ARGUMENT;
    if (n_jobs) *n_jobs = (int)b->jobs.size();
    if (n_groups) {
        std::map<std::pair<Module*, std::string>, int> g;
        for (auto& j : b->jobs) g[std::make_pair(j.m, j.kernel)] = 1;
        *n_groups = (int)g.size();
    }
    if (block) *block = b->plan.block;
    if (waves_per_simd) *waves_per_simd = b->plan.waves_per_simd;
    return UC_OK;
}

}  // extern "C"

// =====================================================================
// Serving: continuous batching. A scheduler thread admits submitted jobs
// in arrival order while their output buffers fit in the HBM budget and
// keeps up to kSlots rounds in flight, one per HIP stream. A round takes
// cells from the admitted jobs (a fair share for every job first, then
// oldest first) up to the GPU's resident lanes, and launches one persistent
// kernel per (module, kernel) group over the round's job table; the next
// round is built and launched as soon as a slot is free, so rounds queue
// back to back on the device and the tail of one overlaps the next. Jobs
// that arrive meanwhile join the next round. Large jobs run in slices
// (sliced kernels, state in a slab): a round advances their cells by about
// slice_steps steps and hands them back, so a round lasts about as long as
// one slice and short jobs never wait behind a long job's cells. Entries of
// small grids run at a higher wave issue priority. A job's cells compute exactly as in a direct launch (same kernel
// code per cell; the output plane stride stays the job's own), so results
// do not depend on what else shares a round. Outputs stay in device memory
// until fetched. Device buffers use the stream-ordered allocator where the
// device has one (hipFree synchronises the whole device).
struct uc_server {
    static constexpr int kSlots = 6;
    static constexpr int kLongClass = 4;   // cells of >= 1e4 steps: at most kLongSlots slots
    static constexpr int kLongSlots = 3;
    static constexpr int kMaxGroups = 64;  // counters per slot
    uc_engine* e = nullptr;
    uc_server_config cfg{};
    struct Job {
        long long id = 0;
        int kind = 0;  // 0 bif, 1 lle, 2 ls, 3 basins
        Module* m = nullptr;
        std::string kernel;
        UcJob tmpl{};
        double* tables = nullptr;
        // owned device outputs
        int* flags = nullptr;
        double* out0 = nullptr;
        double* out1 = nullptr;
        size_t out0_n = 0, out1_n = 0;
        size_t bytes = 0;
        long long cells = 0;
        long long next = 0;        // cells handed to rounds (first units)
        long long done_cells = 0;  // cells whose last units are done
        // resumable slices: the sliced kernel, its slab, the units per round
        // and the cell ranges waiting for their next units
        SliceInfo si;
        long long ups = 0;         // units per round; 0: whole runs
        double* slab = nullptr;
        size_t slab_reserved = 0;  // slab bytes counted at submission
        struct Range { long long first, cells, unit; };
        std::deque<Range> resume;
        int inflight = 0;          // rounds in flight that hold cells of this job
        double step_cost = 1.0;    // traj-steps per cell
        int state = 0;             // 0 queued, 1 running, 2 done, 3 cancelled, 4 failed, 5 fetched
        std::chrono::steady_clock::time_point t_submit, t_done;
        // request copies (pointers inside refer to these)
        std::vector<double> ic, vals;
        uc_common c{};
        uc_bif_request bif{};
        uc_lyap_request lyap{};
        int n_exp = 1, max_peaks = 0, n_vars = 0;
        long long iters = 0;
        // basins: out0 = canonical points [lbl][cells], flags = periods,
        // labels and step counters [2][cells]; the catalogue is built when
        // the last cell is done
        uc_basins_request bas{};
        int* labels = nullptr;
        long long* steps = nullptr;
        int lbl = 0;
        BasinsCatalogue cat;
        std::string error;
    };
    struct Entry {
        std::shared_ptr<Job> job;
        long long first, cells;
        long long u0 = 0, u1 = 0;  // units of this round (sliced jobs)
    };
    struct Slot {
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        UcJob* d_jobs = nullptr;
        long long* d_start = nullptr;
        unsigned long long* d_ctr = nullptr;
        UcJob* h_jobs = nullptr;      // pinned staging, read by the copy until `done`
        long long* h_start = nullptr;
        size_t cap = 0;               // entries (jobs); starts hold cap + kMaxGroups
        bool busy = false;
        int cls = -1;                 // duration class of the round in flight
        std::vector<Entry> entries;
    };
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::shared_ptr<Job>> pending;
    std::vector<std::shared_ptr<Job>> active;
    std::map<long long, std::shared_ptr<Job>> all;
    long long next_id = 1;
    size_t used_bytes = 0, budget = 0;
    double slice_steps = 0.0;  // traj-steps per cell per round of a sliced job; 0: whole runs
    // jobs of more cells than this run in slices; fewer cells hold few wave
    // slots, so they run whole (a thin job sliced would wait for a round per
    // slice)
    long long slice_min_cells = 65536;
    bool stop = false;
    bool async_alloc = false;
    hipMemPool_t pool = nullptr;   // the server's own pool; keeps up to the budget mapped
    std::thread worker;
    hipStream_t stream = nullptr;  // admission uploads, fetch copies, frees
    Slot slots[kSlots];
    int n_busy = 0;
    std::chrono::steady_clock::time_point busy_since;
    std::vector<double> latencies_ms;
    double busy_s = 0.0, traj_steps_done = 0.0;
    long long rounds = 0;
    int last_class = -1;  // duration class of the last round picked
    std::chrono::steady_clock::time_point t_start;
    std::string error;
};

namespace {

hipError_t server_malloc(uc_server* S, void** p, size_t bytes, hipStream_t s) {
    return S->async_alloc ? hipMallocFromPoolAsync(p, bytes, S->pool, s) : hipMalloc(p, bytes);
}

void server_release(uc_server* S, void* p, hipStream_t s) {
    if (p) (void)(S->async_alloc ? hipFreeAsync(p, s) : hipFree(p));
}

// Frees a job's device buffers; no round in flight may hold its cells.
void server_free(uc_server* S, uc_server::Job& j) {
    server_release(S, j.slab, S->stream);
    j.slab = nullptr;
    server_release(S, j.flags, S->stream);
    server_release(S, j.out0, S->stream);
    server_release(S, j.out1, S->stream);
    server_release(S, j.tables, S->stream);
    server_release(S, j.labels, S->stream);
    server_release(S, j.steps, S->stream);
    j.flags = nullptr; j.out0 = j.out1 = nullptr; j.tables = nullptr; j.labels = nullptr; j.steps = nullptr;
}

// admission: allocate outputs + tables, build the descriptor template
bool server_admit(uc_server* S, uc_server::Job& j) {
    const size_t need = table_doubles(j.c);
    std::vector<double> host(need);
    fill_tables(j.n_vars, j.c, host.data());
    hipError_t he = server_malloc(S, (void**)&j.tables, need * sizeof(double), S->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(j.tables, host.data(), need * sizeof(double), hipMemcpyHostToDevice, S->stream);
    if (he == hipSuccess) he = server_malloc(S, (void**)&j.flags, sizeof(int) * (size_t)j.cells, S->stream);
    if (he == hipSuccess) he = server_malloc(S, (void**)&j.out0, sizeof(double) * std::max<size_t>(1, j.out0_n), S->stream);
    if (he == hipSuccess && j.out1_n) he = server_malloc(S, (void**)&j.out1, sizeof(double) * j.out1_n, S->stream);
    if (he == hipSuccess && j.kind == 3) he = server_malloc(S, (void**)&j.labels, sizeof(int) * (size_t)j.cells, S->stream);
    if (he == hipSuccess && j.kind == 3)
        he = server_malloc(S, (void**)&j.steps, 2 * sizeof(long long) * (size_t)j.cells, S->stream);
    if (he != hipSuccess) { (void)hipStreamSynchronize(S->stream); server_free(S, j); j.error = hip_msg(he, "admission"); return false; }
    wire_job(j.c, j.tables, j.tmpl);
    if (j.kind == 3) {
        fill_basins(&j.bas, j.lbl, j.tmpl);
        j.tmpl.flag = j.flags;
        j.tmpl.out0 = j.out0;
        j.tmpl.counters = j.steps;
        j.tmpl.label = j.labels;
    } else if (j.kind == 0) {
        uc_bif_output o{j.flags, j.out0, j.out1, nullptr, 0, nullptr};
        fill_bif(&j.bif, &o, j.iters, j.tmpl);
    } else {
        uc_lyap_output o{j.out0, j.flags, nullptr};
        fill_lyap(&j.lyap, &o, j.tmpl);
    }
    // resumable slices when the job is longer than one round's steps (its
    // slab was counted in j.bytes at submission; given back otherwise)
    j.si = slice_info(j.kind, j.tmpl, j.n_vars, j.step_cost);
    const size_t slab_bytes = sizeof(double) * (size_t)j.si.words * (size_t)j.cells;
    if (S->slice_steps > 0.0 && j.cells > S->slice_min_cells && j.si.units > 0 &&
        slice_units(j.si, S->slice_steps) < j.si.units) {
        j.ups = slice_units(j.si, S->slice_steps);
        j.kernel = sliced_kernel(j.kernel);
        he = server_malloc(S, (void**)&j.slab, slab_bytes, S->stream);
    } else if (j.slab_reserved) {
        j.bytes -= j.slab_reserved;
    }
    j.slab_reserved = 0;
    if (he == hipSuccess) he = hipStreamSynchronize(S->stream);  // tables uploaded, buffers usable on any stream
    if (he != hipSuccess) { server_free(S, j); j.error = hip_msg(he, "admission"); return false; }
    return true;
}

bool slot_reserve(uc_server* S, uc_server::Slot& sl, size_t n_entries) {
    if (n_entries <= sl.cap) return true;
    size_t cap = std::max<size_t>(256, sl.cap);
    while (cap < n_entries) cap *= 2;
    server_release(S, sl.d_jobs, sl.stream);
    server_release(S, sl.d_start, sl.stream);
    if (sl.h_jobs) (void)hipHostFree(sl.h_jobs);
    if (sl.h_start) (void)hipHostFree(sl.h_start);
    sl.d_jobs = nullptr; sl.d_start = nullptr; sl.h_jobs = nullptr; sl.h_start = nullptr; sl.cap = 0;
    const size_t n_start = cap + uc_server::kMaxGroups;
    if (server_malloc(S, (void**)&sl.d_jobs, sizeof(UcJob) * cap, sl.stream) != hipSuccess ||
        server_malloc(S, (void**)&sl.d_start, sizeof(long long) * n_start, sl.stream) != hipSuccess ||
        hipHostMalloc((void**)&sl.h_jobs, sizeof(UcJob) * cap, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&sl.h_start, sizeof(long long) * n_start, hipHostMallocDefault) != hipSuccess)
        return false;
    sl.cap = cap;
    return true;
}

// Traj-steps per cell of one round of the job: the whole run, or one slice.
double round_cost(const uc_server::Job& J) {
    return J.ups ? std::min(J.step_cost, (double)J.ups * J.si.unit_cost) : J.step_cost;
}
// Duration class of a job's rounds: the decade of their per-cell step count.
int server_class(const uc_server::Job& J) { return (int)std::floor(std::log10(std::max(1.0, round_cost(J)))); }
bool server_has_work(const uc_server::Job& J) { return J.state == 1 && (J.next < J.cells || !J.resume.empty()); }

// Picks the cells of the next round (under S->mu) and returns its duration
// class in *cls_out. A round holds cells of one duration class only, so it
// lasts about as long as its cells do and a short job does not wait for
// another job's long cells. The shortest class with cells left goes first
// when no round of it is in flight; otherwise the classes take turns (the
// next class after the last round's, cyclically). Classes of long cells
// (>= kLongClass) are skipped when `allow_long` is false, which keeps a slot
// free for short work. Inside the class every job gets a fair share of the
// slot's lanes first, then the oldest jobs fill the rest; an optional
// traj-step cap bounds the round's work.
std::vector<uc_server::Entry> server_pick(uc_server* S, long long lanes, int block, bool allow_long, int* cls_out) {
    std::vector<uc_server::Entry> out;
    std::vector<uc_server::Job*> open;
    int cls = -1, lowest = -1;
    for (auto& jp : S->active) {
        if (!server_has_work(*jp)) continue;
        const int c = server_class(*jp);
        if (!allow_long && c >= uc_server::kLongClass) continue;
        if (lowest < 0 || c < lowest) lowest = c;
        if (c > S->last_class && (cls < 0 || c < cls)) cls = c;
    }
    if (lowest < 0) return out;
    bool lowest_running = false;
    for (auto& sl : S->slots)
        if (sl.busy && sl.cls == lowest) lowest_running = true;
    if (cls < 0 || !lowest_running) cls = lowest;
    S->last_class = cls;
    *cls_out = cls;
    for (auto& jp : S->active)
        if (server_has_work(*jp) && server_class(*jp) == cls) open.push_back(jp.get());
    const double cap_steps = S->cfg.round_traj_steps > 0 ? S->cfg.round_traj_steps : 0.0;
    std::map<uc_server::Job*, long long> take;
    std::map<std::pair<Module*, std::string>, int> groups;
    long long cells = 0;
    double steps = 0.0;
    const long long share = std::max<long long>(block, (lanes / (long long)open.size()) / block * block);
    auto available = [](const uc_server::Job* J) {
        long long n = J->cells - J->next;
        for (auto& r : J->resume) n += r.cells;
        return n;
    };
    for (int pass = 0; pass < 2 && cells < lanes; ++pass) {
        for (uc_server::Job* J : open) {
            if (cells >= lanes || (cap_steps > 0.0 && steps >= cap_steps)) break;
            const auto g = std::make_pair(J->m, J->kernel);
            if (!groups.count(g) && (int)groups.size() >= uc_server::kMaxGroups) continue;
            const long long left = available(J) - take[J];
            const double cost = round_cost(*J);
            long long n = std::min(left, pass == 0 ? share : lanes - cells);
            if (cap_steps > 0.0) n = std::min<long long>(n, std::max<long long>(1, (long long)((cap_steps - steps) / cost)));
            if (n <= 0) continue;
            take[J] += n;
            cells += n;
            steps += (double)n * cost;
            groups[g] = 1;
        }
    }
    // cells waiting for their next units first, then new cells
    for (auto& jp : S->active) {
        auto it = take.find(jp.get());
        if (it == take.end() || it->second <= 0) continue;
        uc_server::Job& J = *jp;
        long long n = it->second;
        while (n > 0 && !J.resume.empty()) {
            auto& r = J.resume.front();
            const long long k = std::min(n, r.cells);
            out.push_back({jp, r.first, k, r.unit, std::min(J.si.units, r.unit + J.ups)});
            r.first += k;
            r.cells -= k;
            n -= k;
            if (r.cells == 0) J.resume.pop_front();
            ++J.inflight;
        }
        if (n > 0) {
            out.push_back({jp, J.next, n, 0, J.ups ? std::min(J.si.units, J.ups) : 0});
            J.next += n;
            ++J.inflight;
        }
    }
    return out;
}

// Uploads the round's job tables and launches it on the slot's stream.
// Returns an error message, empty on success.
std::string server_launch(uc_server* S, uc_server::Slot& sl, int block, int wps, int cus) {
    // group entries by (module, kernel), longest cells first inside a group
    std::map<std::pair<Module*, std::string>, std::vector<const uc_server::Entry*>> groups;
    for (auto& en : sl.entries) groups[std::make_pair(en.job->m, en.job->kernel)].push_back(&en);
    if (!slot_reserve(S, sl, sl.entries.size())) return "server: job-table buffers";
    size_t jo = 0, so = 0;
    struct G { Module* m; const std::string* k; size_t jofs, sofs; int nj; long long chunks; };
    std::vector<G> gl;
    for (auto& kv : groups) {
        auto& v = kv.second;
        std::stable_sort(v.begin(), v.end(), [](const uc_server::Entry* a, const uc_server::Entry* b) {
            return a->job->step_cost > b->job->step_cost;
        });
        G g{kv.first.first, &kv.first.second, jo, so, (int)v.size(), 0};
        sl.h_start[so++] = 0;
        for (const uc_server::Entry* en : v) {
            UcJob j = en->job->tmpl;
            j.cell_offset = en->first;
            j.n_cells = en->cells;
            j.stride = en->job->cells;
            j.flag = en->job->flags + en->first;
            j.out0 = en->job->out0 + en->first;
            j.out1 = en->job->out1 ? en->job->out1 + en->first : nullptr;
            j.counters = en->job->steps ? en->job->steps + en->first : nullptr;
            if (en->job->ups) {
                j.slab = en->job->slab + en->first;
                j.unit_begin = en->u0;
                j.unit_end = en->u1;
            }
            // issue priority by job size: small grids cannot fill the GPU
            // and their latency is the pace of their own few waves
            const long long nc = en->job->cells;
            j.priority = nc <= 16384 ? 3 : nc <= 65536 ? 2 : nc <= 262144 ? 1 : 0;
            sl.h_jobs[jo++] = j;
            g.chunks += (en->cells + block - 1) / block;
            sl.h_start[so++] = g.chunks;
        }
        gl.push_back(g);
    }
    hipError_t he = hipMemcpyAsync(sl.d_jobs, sl.h_jobs, sizeof(UcJob) * jo, hipMemcpyHostToDevice, sl.stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(sl.d_start, sl.h_start, sizeof(long long) * so, hipMemcpyHostToDevice, sl.stream);
    hipFunction_t f;
    if (he == hipSuccess) he = hipModuleGetFunction(&f, gl[0].m->mod, "uc_batch_reset");
    int ng = (int)gl.size();
    void* rargs[] = {&sl.d_ctr, &ng};
    if (he == hipSuccess) he = hipModuleLaunchKernel(f, 1, 1, 1, 64, 1, 1, 0, sl.stream, rargs, nullptr);
    const long long resident = (long long)cus * 4 * wps / (block / 64);
    for (size_t gi = 0; gi < gl.size() && he == hipSuccess; ++gi) {
        G& g = gl[gi];
        he = hipModuleGetFunction(&f, g.m->mod, g.k->c_str());
        UcJob* dj = sl.d_jobs + g.jofs;
        long long* ds = sl.d_start + g.sofs;
        unsigned long long* ctr = sl.d_ctr + gi;
        void* args[] = {&dj, &ds, &g.nj, &ctr};
        const long long grid = std::min<long long>(resident, std::max<long long>(1, g.chunks));
        if (he == hipSuccess)
            he = hipModuleLaunchKernel(f, (unsigned)grid, 1, 1, (unsigned)block, 1, 1, 0, sl.stream, args, nullptr);
    }
    if (he == hipSuccess) he = hipEventRecord(sl.done, sl.stream);
    return he == hipSuccess ? std::string() : hip_msg(he, "server round");
}

// Books a finished (or failed) round (under S->mu).
void server_retire(uc_server* S, uc_server::Slot& sl, const std::string& why) {
    const auto now = std::chrono::steady_clock::now();
    for (auto& en : sl.entries) {
        auto& J = *en.job;
        --J.inflight;
        if (!why.empty()) {
            if (J.state == 1) { J.state = 4; J.error = why; J.t_done = now; }
            continue;
        }
        // fixed-step work only
        if (J.kind != 3)
            S->traj_steps_done += J.ups ? (double)en.cells * (double)(en.u1 - en.u0) * J.si.unit_cost
                                        : (double)en.cells * J.step_cost;
        if (J.ups && en.u1 < J.si.units) {  // more units to run: back in the queue
            if (J.state == 1) J.resume.push_back({en.first, en.cells, en.u1});
            continue;
        }
        J.done_cells += en.cells;
        if (J.state == 1 && J.done_cells >= J.cells) {
            if (J.kind == 3) {
                UcJob g = J.tmpl;  // whole grid
                g.cell_offset = 0;
                g.n_cells = J.cells;
                g.stride = J.cells;
                const std::string err = basins_catalogue(
                    J.m, g, J.bas.quantum, S->stream,
                    [S](void** ptr, size_t b) { return server_malloc(S, ptr, b, S->stream); },
                    [S](void* ptr) { server_release(S, ptr, S->stream); }, J.cat);
                if (!err.empty()) { J.state = 4; J.error = err; J.t_done = std::chrono::steady_clock::now(); continue; }
            }
            J.state = 2;
            J.t_done = std::chrono::steady_clock::now();
            S->latencies_ms.push_back(std::chrono::duration<double, std::milli>(J.t_done - J.t_submit).count());
        }
    }
    sl.entries.clear();
    sl.busy = false;
    if (--S->n_busy == 0) S->busy_s += std::chrono::duration<double>(now - S->busy_since).count();
    ++S->rounds;
}

// Drops finished jobs from the active list and frees cancelled / failed
// ones once no round holds their cells (under S->mu).
void server_sweep(uc_server* S) {
    std::vector<std::shared_ptr<uc_server::Job>> still;
    for (auto& jp : S->active) {
        if (jp->state == 1 || jp->inflight > 0) { still.push_back(jp); continue; }
        if (jp->state != 2) { S->used_bytes -= jp->bytes; server_free(S, *jp); }
    }
    S->active.swap(still);
}

void server_destroy_slots(uc_server* S) {
    for (auto& sl : S->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        server_release(S, sl.d_jobs, sl.stream);
        server_release(S, sl.d_start, sl.stream);
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.d_ctr) (void)hipFree(sl.d_ctr);
        if (sl.h_jobs) (void)hipHostFree(sl.h_jobs);
        if (sl.h_start) (void)hipHostFree(sl.h_start);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        sl = uc_server::Slot();
    }
}

void server_loop(uc_server* S) {
    (void)hipSetDevice(S->e->device);
    const int block = S->cfg.block > 0 ? S->cfg.block : 256;
    const int wps = S->cfg.waves_per_simd > 0 ? S->cfg.waves_per_simd : 8;
    const int cus = cu_count(S->e->device);
    // a round holds the GPU's resident lanes: the slots' streams share fewer
    // hardware queues than there are slots, so each round must fill the GPU
    // on its own; the rounds queued behind it start as soon as it drains
    const long long lanes = std::max<long long>(block, (long long)cus * 4 * wps * 64 / block * block);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        // retire finished rounds
        for (auto& sl : S->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            server_retire(S, sl, q == hipSuccess ? std::string() : hip_msg(q, "server round"));
        }
        if (S->stop)
            for (auto& jp : S->active)
                if (jp->state == 1) { jp->state = 3; jp->t_done = std::chrono::steady_clock::now(); }
        server_sweep(S);
        S->cv.notify_all();
        // admission in arrival order while the outputs fit
        while (!S->stop && !S->pending.empty()) {
            auto j = S->pending.front();
            if (j->state == 3) { S->pending.pop_front(); continue; }
            // over the budget a job waits, unless no admitted job still holds
            // memory (done but unfetched jobs count): a job larger than the
            // whole budget then runs alone
            if (S->used_bytes + j->bytes > S->budget && S->used_bytes > 0) break;
            S->pending.pop_front();
            if (!server_admit(S, *j)) { j->state = 4; j->t_done = std::chrono::steady_clock::now(); continue; }
            S->used_bytes += j->bytes;
            j->state = 1;
            S->active.push_back(j);
        }
        // launch on a free slot
        uc_server::Slot* free_slot = nullptr;
        for (auto& sl : S->slots)
            if (!sl.busy && sl.stream) { free_slot = &sl; break; }
        bool launched = false;
        if (free_slot && !S->stop) {
            int long_busy = 0, cls = -1;
            for (auto& sl : S->slots)
                if (sl.busy && sl.cls >= uc_server::kLongClass) ++long_busy;
            auto entries = server_pick(S, lanes, block, long_busy < uc_server::kLongSlots, &cls);
            if (!entries.empty()) {
                free_slot->entries = std::move(entries);
                free_slot->busy = true;
                free_slot->cls = cls;
                if (S->n_busy++ == 0) S->busy_since = std::chrono::steady_clock::now();
                const std::string why = server_launch(S, *free_slot, block, wps, cus);
                if (!why.empty()) {
                    (void)hipStreamSynchronize(free_slot->stream);
                    server_retire(S, *free_slot, why);
                }
                launched = true;
            }
        }
        if (launched) continue;
        if (S->n_busy == 0) {
            if (S->stop) break;
            if (S->active.empty() || S->pending.empty())
                S->cv.wait_for(lk, std::chrono::milliseconds(20));
            continue;
        }
        // rounds in flight and nothing to launch: poll again shortly (a
        // submission wakes the loop at once)
        S->cv.wait_for(lk, std::chrono::microseconds(200));
    }
}

int server_submit(uc_server* S, std::shared_ptr<uc_server::Job> j, long long* id) {
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) return S->e->fail(UC_ERR_ARGUMENT, "server stopped");
        j->id = S->next_id++;
        j->t_submit = std::chrono::steady_clock::now();
        S->all[j->id] = j;
        S->pending.push_back(j);
    }
    S->cv.notify_all();
    if (id) *id = j->id;
    return UC_OK;
}

// Counts the slab of a job that may run in slices (no sweep over h) in its
// bytes; admission gives the bytes back when the job runs whole.
void server_reserve_slab(uc_server* S, uc_server::Job& j) {
    bool step_axis = false;
    for (int i = 0; i < j.c.n_axes; ++i) step_axis |= j.c.axis[i].kind == UC_AXIS_STEP;
    // sliced exactly when one round's units are fewer than the job's, i.e.
    // when the job's steps per cell exceed the round's
    if (S->slice_steps <= 0.0 || step_axis || j.step_cost <= S->slice_steps || j.cells <= S->slice_min_cells) return;
    const int n = j.n_vars;
    const int words = j.kind == kKindBif ? UC_SLAB_BIF_WORDS(n)
                      : j.kind == kKindBasins ? UC_SLAB_BAS_WORDS(n)
                                              : UC_SLAB_LYAP_WORDS(n, j.kind == kKindLs ? n : 1);
    j.slab_reserved = sizeof(double) * (size_t)words * (size_t)j.cells;
    j.bytes += j.slab_reserved;
}

// copy the request so the job does not depend on caller memory
void server_copy_common(uc_server::Job& j, const uc_common& c, int n_vars) {
    j.c = c;
    j.n_vars = n_vars;
    j.ic.assign(c.initial_conditions ? c.initial_conditions : nullptr,
                c.initial_conditions ? c.initial_conditions + n_vars : nullptr);
    j.vals.assign(c.base_values, c.base_values + c.n_values);
    j.c.initial_conditions = j.ic.empty() ? nullptr : j.ic.data();
    j.c.base_values = j.vals.data();
    j.cells = cells_of(c);
}

}  // namespace

extern "C" {

int uc_server_create(uc_engine* e, const uc_server_config* cfg, uc_server** out) {
    if (!e || !out) return UC_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    int rc = enter(e);
    if (rc) return rc;
    auto S = new uc_server;
    S->e = e;
    if (cfg) S->cfg = *cfg;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    const double frac = S->cfg.hbm_fraction > 0.0 ? S->cfg.hbm_fraction : 0.5;
    S->budget = S->cfg.max_bytes > 0 ? (size_t)S->cfg.max_bytes : (size_t)(frac * (double)free_b);
    S->slice_steps = S->cfg.slice_steps == 0.0 ? 1024.0 : (S->cfg.slice_steps > 0.0 ? S->cfg.slice_steps : 0.0);
    S->slice_min_cells = S->cfg.slice_min_cells == 0 ? 65536 : std::max(0LL, S->cfg.slice_min_cells);
    int pools = 0;
    S->async_alloc = hipDeviceGetAttribute(&pools, hipDeviceAttributeMemoryPoolsSupported, e->device) == hipSuccess &&
                     pools != 0;
    if (S->async_alloc) {
        // A pool of its own whose release threshold is the budget: memory a
        // finished job frees stays mapped for the next job instead of going
        // back to the driver at every synchronisation (the default pool's
        // threshold is 0, and re-mapping stalls admissions under load).
        hipMemPoolProps props{};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = e->device;
        uint64_t keep = (uint64_t)S->budget;
        if (hipMemPoolCreate(&S->pool, &props) != hipSuccess ||
            hipMemPoolSetAttribute(S->pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) {
            if (S->pool) (void)hipMemPoolDestroy(S->pool);
            S->pool = nullptr;
            S->async_alloc = false;
        }
    }
    hipError_t he = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
    for (auto& sl : S->slots) {
        if (he == hipSuccess) he = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (he == hipSuccess) he = hipMalloc((void**)&sl.d_ctr, sizeof(unsigned long long) * uc_server::kMaxGroups);
    }
    if (he != hipSuccess) {
        server_destroy_slots(S);
        if (S->stream) (void)hipStreamDestroy(S->stream);
        if (S->pool) (void)hipMemPoolDestroy(S->pool);
        delete S;
        return e->fail(UC_ERR_HIP, hip_msg(he, "server streams"));
    }
    S->t_start = std::chrono::steady_clock::now();
    e->refs.fetch_add(1, std::memory_order_relaxed);
    S->worker = std::thread(server_loop, S);
    *out = S;
    return UC_OK;
}

void uc_server_destroy(uc_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        for (auto& j : S->pending) j->state = 3;
    }
    S->cv.notify_all();
    if (S->worker.joinable()) S->worker.join();
    uc_engine* e = S->e;
    {
        std::lock_guard<std::recursive_mutex> lk(e->mu);
        (void)hipSetDevice(e->device);
        for (auto& kv : S->all) {
            if (kv.second->m) --kv.second->m->pins;
            server_free(S, *kv.second);
        }
        server_destroy_slots(S);
        if (S->stream) { (void)hipStreamSynchronize(S->stream); (void)hipStreamDestroy(S->stream); }
        if (S->pool) (void)hipMemPoolDestroy(S->pool);  // after every free above has completed
        delete S;
    }
    engine_release(e);
}

int uc_server_submit_lyap(uc_server* S, uc_system* s, const uc_lyap_request* r, int spectrum, long long* id) {
    if (!S || !s || !r) return UC_ERR_ARGUMENT;
    uc_engine* e = S->e;
    precompile(e, s, r->c.strict != 0);
    auto j = std::make_shared<uc_server::Job>();
    {
        std::lock_guard<std::recursive_mutex> lk(e->mu);
        double dummy = 0.0;
        int dflag = 0;
        uc_lyap_output probe{&dummy, &dflag, nullptr};
        int rc = prep_lyap(e, s, r, &probe);
        if (rc) return rc;
        if ((rc = whole_grid_only(e, r->c, "a served job"))) return rc;
        if ((rc = enter(e))) return rc;
        if ((rc = get_module(e, s, r->c.strict != 0, true, &j->m))) return rc;
        ++j->m->pins;
    }
    server_copy_common(*j, r->c, s->n_vars);
    j->lyap = *r;
    j->lyap.c = j->c;
    j->kind = spectrum ? 2 : 1;
    j->n_exp = spectrum ? s->n_vars : 1;
    j->kernel = std::string(lyap_kernel(spectrum != 0, r->c.precision)) + "_q";
    j->out0_n = (size_t)j->n_exp * (size_t)j->cells;
    j->bytes = j->out0_n * sizeof(double) + (size_t)j->cells * sizeof(int);
    const long long skip = r->c.h > 0 ? (long long)(r->c.transient_time / r->c.h) : 0;
    const long long blocks = (long long)(r->c.t_max / r->NT), nts = r->c.h > 0 ? (long long)(r->NT / r->c.h) : 1;
    j->step_cost = (double)(skip + 1 + blocks * (nts + 1));
    server_reserve_slab(S, *j);
    return server_submit(S, j, id);
}

int uc_server_submit_bif(uc_server* S, uc_system* s, const uc_bif_request* r, long long* id) {
    if (!S || !s || !r) return UC_ERR_ARGUMENT;
    uc_engine* e = S->e;
    precompile(e, s, r->c.strict != 0);
    auto j = std::make_shared<uc_server::Job>();
    long long iters = 0;
    {
        std::lock_guard<std::recursive_mutex> lk(e->mu);
        int dflag = 0;
        double dummy = 0.0;
        uc_bif_output probe{&dflag, &dummy, &dummy, nullptr, 0, nullptr};
        int rc = prep_bif(e, s, r, &probe, &iters);
        if (rc) return rc;
        if ((rc = whole_grid_only(e, r->c, "a served job"))) return rc;
        if ((rc = enter(e))) return rc;
        if ((rc = get_module(e, s, r->c.strict != 0, true, &j->m))) return rc;
        ++j->m->pins;
    }
    server_copy_common(*j, r->c, s->n_vars);
    j->bif = *r;
    j->bif.c = j->c;
    j->kind = 0;
    j->max_peaks = r->max_peaks;
    j->iters = iters;
    j->kernel = r->c.precision == UC_PREC_F64 ? "uc_bif_f64_q" : "uc_bif_f32c_q";
    j->out0_n = j->out1_n = (size_t)std::max(1, r->max_peaks) * (size_t)j->cells;
    j->bytes = 2 * j->out0_n * sizeof(double) + (size_t)j->cells * sizeof(int);
    const long long skip = r->c.h > 0 ? (long long)(r->c.transient_time / r->c.h) : 0;
    j->step_cost = (double)(skip + 1 + iters * r->pre_scaller + r->pre_scaller);
    server_reserve_slab(S, *j);
    return server_submit(S, j, id);
}

int uc_server_submit_basins(uc_server* S, uc_system* s, const uc_basins_request* r, long long* id) {
    if (!S || !s || !r) return UC_ERR_ARGUMENT;
    uc_engine* e = S->e;
    precompile(e, s, r->c.strict != 0);
    auto j = std::make_shared<uc_server::Job>();
    uc_common c;
    {
        std::lock_guard<std::recursive_mutex> lk(e->mu);
        int rc = prep_basins(e, s, r, c, j->lbl);
        if (rc) return rc;
        if ((rc = enter(e))) return rc;
        if ((rc = get_module(e, s, c.strict != 0, true, &j->m))) return rc;
        ++j->m->pins;
    }
    server_copy_common(*j, c, s->n_vars);
    j->bas = *r;
    j->bas.c = j->c;
    j->kind = 3;
    j->kernel = c.precision == UC_PREC_F64 ? "uc_bas_f64_q" : "uc_bas_f32c_q";
    j->out0_n = (size_t)j->lbl * (size_t)j->cells;
    j->bytes = (size_t)j->cells * (2 * sizeof(int) + 2 * sizeof(long long) + (size_t)j->lbl * sizeof(double));
    // scheduling weight only (the integrator is adaptive)
    j->step_cost = basins_section_cost(c.precision) * (double)r->n_sections;
    server_reserve_slab(S, *j);
    return server_submit(S, j, id);
}

int uc_server_poll(uc_server* S, long long id, int* state, double* latency_ms) {
    if (!S) return UC_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return S->e->fail(UC_ERR_ARGUMENT, "unknown job id");
    auto& j = *it->second;
    if (state) *state = j.state;
    if (latency_ms) *latency_ms = (j.state >= 2) ? std::chrono::duration<double, std::milli>(j.t_done - j.t_submit).count() : -1.0;
    return UC_OK;
}

int uc_server_wait(uc_server* S, long long id, double timeout_s) {
    if (!S) return UC_ERR_ARGUMENT;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return S->e->fail(UC_ERR_ARGUMENT, "unknown job id");
    auto j = it->second;
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
    while (j->state < 2) {
        if (S->cv.wait_until(lk, deadline) == std::cv_status::timeout && j->state < 2) return UC_ERR_BUSY;
    }
    return j->state == 2 ? UC_OK : (j->state == 3 ? UC_ERR_CANCELLED : UC_ERR_HIP);
}

int uc_server_cancel(uc_server* S, long long id) {
    if (!S) return UC_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return S->e->fail(UC_ERR_ARGUMENT, "unknown job id");
    if (it->second->state <= 1) { it->second->state = 3; it->second->t_done = std::chrono::steady_clock::now(); }
    S->cv.notify_all();
    return UC_OK;
}

// Copies a finished job's outputs to host memory and releases them.
// lyap: values[n_exp][cells], flags[cells]; bif: flags, peaks/intervals [max_peaks][cells].
int uc_server_fetch(uc_server* S, long long id, int* flags, double* out0, double* out1) {
    if (!S) return UC_ERR_ARGUMENT;
    std::shared_ptr<uc_server::Job> j;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->all.find(id);
        if (it == S->all.end()) return S->e->fail(UC_ERR_ARGUMENT, "unknown job id");
        j = it->second;
        if (j->state != 2) return j->state < 2 ? UC_ERR_BUSY : S->e->fail(UC_ERR_ARGUMENT, "job did not complete: " + j->error);
        if (j->kind == 3) return S->e->fail(UC_ERR_ARGUMENT, "basins job: use uc_server_fetch_basins");
    }
    (void)hipSetDevice(S->e->device);
    hipError_t he = hipSuccess;
    if (flags) he = hipMemcpyAsync(flags, j->flags, sizeof(int) * (size_t)j->cells, hipMemcpyDeviceToHost, S->stream);
    if (he == hipSuccess && out0)
        he = hipMemcpyAsync(out0, j->out0, sizeof(double) * j->out0_n, hipMemcpyDeviceToHost, S->stream);
    if (he == hipSuccess && out1 && j->out1)
        he = hipMemcpyAsync(out1, j->out1, sizeof(double) * j->out1_n, hipMemcpyDeviceToHost, S->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(S->stream);
    if (he != hipSuccess) return S->e->fail(UC_ERR_HIP, hip_msg(he, "fetch"));
    std::lock_guard<std::mutex> lk(S->mu);
    S->used_bytes -= j->bytes;
    server_free(S, *j);
    if (j->m) { --j->m->pins; j->m = nullptr; }
    j->state = 5;  // fetched
    S->cv.notify_all();
    return UC_OK;
}

// Copies a finished basins job's results to host memory (labels and periods
// required; cycle_point [2][cells], steps [2][cells] and the attractor
// catalogue when given) and releases its device buffers.
int uc_server_fetch_basins(uc_server* S, long long id, const uc_basins_output* o) {
    if (!S || !o || !o->labels || !o->periods) return UC_ERR_ARGUMENT;
    std::shared_ptr<uc_server::Job> j;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->all.find(id);
        if (it == S->all.end()) return S->e->fail(UC_ERR_ARGUMENT, "unknown job id");
        j = it->second;
        if (j->state != 2) return j->state < 2 ? UC_ERR_BUSY : S->e->fail(UC_ERR_ARGUMENT, "job did not complete: " + j->error);
        if (j->kind != 3) return S->e->fail(UC_ERR_ARGUMENT, "not a basins job: use uc_server_fetch");
    }
    (void)hipSetDevice(S->e->device);
    const size_t n = (size_t)j->cells;
    hipError_t he = hipMemcpyAsync(o->labels, j->labels, sizeof(int) * n, hipMemcpyDeviceToHost, S->stream);
    if (he == hipSuccess) he = hipMemcpyAsync(o->periods, j->flags, sizeof(int) * n, hipMemcpyDeviceToHost, S->stream);
    if (he == hipSuccess && o->cycle_point)
        he = hipMemcpyAsync(o->cycle_point, j->out0, sizeof(double) * j->out0_n, hipMemcpyDeviceToHost, S->stream);
    if (he == hipSuccess && o->steps)
        he = hipMemcpyAsync(o->steps, j->steps, 2 * sizeof(long long) * n, hipMemcpyDeviceToHost, S->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(S->stream);
    if (he != hipSuccess) return S->e->fail(UC_ERR_HIP, hip_msg(he, "fetch"));
    copy_catalogue(j->cat, o);
    std::lock_guard<std::mutex> lk(S->mu);
    S->used_bytes -= j->bytes;
    server_free(S, *j);
    if (j->m) { --j->m->pins; j->m = nullptr; }
    j->state = 5;  // fetched
    S->cv.notify_all();
    return UC_OK;
}

int uc_server_stats(uc_server* S, uc_server_stats_t* st) {
    if (!S || !st) return UC_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(S->mu);
    std::vector<double> l = S->latencies_ms;
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) { return l.empty() ? 0.0 : l[(size_t)std::min<double>((double)l.size() - 1, std::floor(q * (double)l.size()))]; };
    st->completed = (long long)l.size();
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    st->max_ms = l.empty() ? 0.0 : l.back();
    const auto now = std::chrono::steady_clock::now();
    st->busy_s = S->busy_s + (S->n_busy > 0 ? std::chrono::duration<double>(now - S->busy_since).count() : 0.0);
    st->wall_s = std::chrono::duration<double>(now - S->t_start).count();
    st->traj_steps = S->traj_steps_done;
    st->rounds = S->rounds;
    st->queued = (long long)S->pending.size();
    st->active = (long long)S->active.size();
    st->used_bytes = (double)S->used_bytes;
    st->budget_bytes = (double)S->budget;
    return UC_OK;
}

}  // extern "C"
