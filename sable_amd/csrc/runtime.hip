// libsable_amd.so: C-ABI runtime of the VBR SpMV/SpMM executor on gfx950.
//
// Responsibilities: device selection, plan upload into HBM, per-plan kernel
// specialisation through hipRTC (with an on-disk code-object cache and the
// ahead-of-time generic kernels as fallback), launches in three scheduling
// modes, HIP-event timing with optional cache flushing,
// graph-captured multi-plan launches, and the router MLP kernels.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <array>
#include <vector>
#include <sys/stat.h>
#include <sys/time.h>
#include <unistd.h>

#include "../../include/sable.h"
#include "plan.h"
#include "sable_kernels.h"
#include "sable_mfma.h"
#include "sable_mlp.h"
#include "sable_partition.h"
#include "sable_inspect.h"

// ---------------------------------------------------------------------------
// ahead-of-time generic instances (fallback when hipRTC is unavailable)

extern "C" __global__ void __launch_bounds__(256, SB_SPMV_WAVES(4)) sb_spmv_aot(SbSpmvArgs a)
{
    sb_spmv_body<256, 4, 4096>(a);
}

extern "C" __global__ void __launch_bounds__(256) sb_spmm_f32_aot(SbSpmmArgs a)
{
    sb_spmm_f32_body<256, 8, 0, 4096>(a);
}

extern "C" __global__ void __launch_bounds__(256, SB_MFMA_WAVES) sb_spmm_mfma_aot(SbSpmmArgs a)
{
    sb_spmm_mfma_body<256, 0, 4096, 3>(a);
}

extern "C" __global__ void __launch_bounds__(256, SB_MFMA_WAVES) sb_spmm_mfma2_aot(SbSpmmArgs a)
{
    sb_spmm_mfma_body<256, 0, 4096, 2>(a);
}

extern "C" __global__ void __launch_bounds__(256) sb_mfma_fixup(const SbTile *fix_tiles, int nfix,
                                                                 const SbMatBufs *mats, const float *part, float *y,
                                                                 int n_rhs)
{
    sb_mfma_fixup_body<256>(fix_tiles, nfix, mats, part, y, n_rhs);
}

extern "C" __global__ void sb_fill_u32(uint32_t *p, int64_t n, uint32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Bandwidth probes: a streaming read (float4 loads, one partial sum stored
// per thread so nothing is optimised away) and a streaming copy.  Each
// workgroup streams its own contiguous chunk (the access pattern of the
// library's tiles; grid-stride loops over the whole buffer measure ~12%
// lower).  They give the attainable HBM rates of the device and the bytes a
// known-size copy shows in the traffic counters.
extern "C" __global__ void __launch_bounds__(256) sb_read_probe(const sb_f4 *__restrict__ p, int64_t n4,
                                                                int64_t chunk4, float *__restrict__ out)
{
    sb_f4 acc = {0.f, 0.f, 0.f, 0.f};
    const int64_t b0 = (int64_t)blockIdx.x * chunk4;
    const int64_t e = b0 + chunk4 < n4 ? b0 + chunk4 : n4;
    int64_t i = b0 + threadIdx.x;
    for (; i + 3 * 256 < e; i += 4 * 256) {
        const sb_f4 a = __builtin_nontemporal_load(&p[i]);
        const sb_f4 b = __builtin_nontemporal_load(&p[i + 256]);
        const sb_f4 c = __builtin_nontemporal_load(&p[i + 512]);
        const sb_f4 d = __builtin_nontemporal_load(&p[i + 768]);
        acc += a + b + c + d;
    }
    for (; i < e; i += 256) acc += __builtin_nontemporal_load(&p[i]);
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc.x + acc.y + acc.z + acc.w;
}

extern "C" __global__ void __launch_bounds__(256) sb_copy_probe(const sb_f4 *__restrict__ src,
                                                                sb_f4 *__restrict__ dst, int64_t n4, int64_t chunk4)
{
    const int64_t b0 = (int64_t)blockIdx.x * chunk4;
    const int64_t e = b0 + chunk4 < n4 ? b0 + chunk4 : n4;
    for (int64_t i = b0 + threadIdx.x; i < e; i += 256)
        __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}

// embedded kernel sources for hipRTC (generated at build time)
#include "_embedded_src.inc"

// Scratch of sable_spmm_noalloc: the partial sums of a plan's split MFMA
// tiles, in a file-scope device array that every GPU the library runs on
// holds until the process ends.  Host code never passes the
// array's name to a kernel: it is a host-side stand-in; its device address
// per GPU comes from hipGetSymbolAddress in sable_init.
#define SABLE_NOALLOC_SCRATCH_BYTES (64LL << 20)
static __device__ __align__(256) unsigned char scratch[SABLE_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;
std::mutex g_mu;
int g_device = -1;
int g_ncu = 0;
void *g_scratch[64] = {};   // device address of `scratch` per GPU (set by sable_init)
int64_t g_hbm = 0;
std::string g_arch = "gfx950";
std::string g_cache_dir;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPOK(expr)                                                                           \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) {                                                               \
            return fail(_e == hipErrorOutOfMemory ? SABLE_ERR_OOM : SABLE_ERR_HIP,            \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                   \
        }                                                                                     \
    } while (0)

uint64_t fnv1a(const std::string &s)
{
    uint64_t h = 1469598103934665603ULL;
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ULL;
    }
    return h;
}

struct RtcKernel {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
};
std::map<std::string, RtcKernel> g_rtc;   // key -> loaded kernel

std::string cache_dir()
{
    if (!g_cache_dir.empty()) return g_cache_dir;
    const char *e = getenv("SABLE_CACHE_DIR");
    if (e && *e) return e;
    return "/tmp/sable_amd_rtc";
}

void mkdirs(const std::string &d)
{
    std::string cur;
    std::stringstream ss(d);
    std::string part;
    if (!d.empty() && d[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        mkdir(cur.c_str(), 0755);
    }
}

// Compiles `entry` (a kernel defined in the generated stub) with the given
// macro definitions; returns the code object.  No device needed.
int rtc_compile(const std::string &defines, const std::string &stub, std::vector<char> &code,
                std::string &log)
{
    const std::string src = defines + "#include \"sable_kernels.h\"\n#include \"sable_mfma.h\"\n" + stub;
    const char *hdrs[] = {k_src_sable_device_h, k_src_sable_kernels_h, k_src_sable_mfma_h};
    const char *names[] = {"sable_device.h", "sable_kernels.h", "sable_mfma.h"};
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "sable_rtc.hip", 3, hdrs, names) != HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return SABLE_ERR_RTC;
    }
    const std::string arch = "--offload-arch=" + g_arch;
    const char *opts[] = {arch.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature",
                          "-Xclang", "-packed-fp32-ops"};
    hiprtcResult r = hiprtcCompileProgram(prog, 7, opts);
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    log.assign(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return SABLE_ERR_RTC;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return SABLE_OK;
}

// Returns (compiling and caching if needed) the code object for a key.
int rtc_code_object(const std::string &defines, const std::string &stub, std::vector<char> &code,
                    std::string &path_out)
{
    int vmaj = 0, vmin = 0;
    hiprtcVersion(&vmaj, &vmin);
    const std::string key = g_arch + "\n" + std::to_string(vmaj) + "." + std::to_string(vmin) + "\n" + defines +
                            stub + k_src_sable_kernels_h + k_src_sable_mfma_h + k_src_sable_device_h;
    char name[64];
    snprintf(name, sizeof name, "sb_%016llx.co", (unsigned long long)fnv1a(key));
    const std::string dir = cache_dir();
    const std::string path = dir + "/" + name;
    path_out = path;
    {
        std::ifstream f(path, std::ios::binary);
        if (f) {
            code.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
            if (!code.empty()) {
                utimes(path.c_str(), nullptr);   // marks the entry as in use (cache pruning)
                return SABLE_OK;
            }
        }
    }
    std::string log;
    int rc = rtc_compile(defines, stub, code, log);
    if (rc) return fail(rc, "hipRTC compile failed: " + log.substr(0, 2000));
    mkdirs(dir);
    const std::string tmp = path + ".tmp" + std::to_string(getpid());
    {
        std::ofstream o(tmp, std::ios::binary);
        o.write(code.data(), (std::streamsize)code.size());
    }
    rename(tmp.c_str(), path.c_str());
    return SABLE_OK;
}

int rtc_get(const std::string &defines, const std::string &stub, const char *entry, hipFunction_t *fn)
{
    std::lock_guard<std::mutex> lk(g_mu);
    const std::string key = defines + stub;
    auto it = g_rtc.find(key);
    if (it != g_rtc.end()) {
        *fn = it->second.fn;
        return SABLE_OK;
    }
    std::vector<char> code;
    std::string path;
    int rc = rtc_code_object(defines, stub, code, path);
    if (rc) return rc;
    RtcKernel k;
    HIPOK(hipModuleLoadData(&k.mod, code.data()));
    HIPOK(hipModuleGetFunction(&k.fn, k.mod, entry));
    g_rtc[key] = k;
    *fn = k.fn;
    return SABLE_OK;
}

// ---------------------------------------------------------------------------
// specialisations

struct Spec {
    int wg = 256, unroll = 4, xcap = 4096, nrhs = 0, rtmax = 8, split = 3;
};

// The macro block of a specialisation.  SABLE_RTC_DEFINES (e.g.
// "SB_MFMA_DEPTH=3;SB_MFMA_XDEPTH=1") adds kernel tuning macros for
// experiments; they are part of the cache key like the rest.
std::string spec_defines(const Spec &s)
{
    char b[320];
    snprintf(b, sizeof b,
             "#define SB_WG %d\n#define SB_UNROLL %d\n#define SB_XCAP %d\n#define SB_NRHS %d\n#define SB_RTMAX %d\n"
             "#define SB_SPLIT %d\n",
             s.wg, s.unroll, s.xcap, s.nrhs, s.rtmax, s.split);
    std::string d = b;
    if (const char *extra = getenv("SABLE_RTC_DEFINES")) {
        std::string e = extra;
        size_t i = 0;
        while (i < e.size()) {
            size_t j = e.find(';', i);
            if (j == std::string::npos) j = e.size();
            std::string item = e.substr(i, j - i);
            const size_t eq = item.find('=');
            if (!item.empty() && item.find_first_not_of("ABCDEFGHIJKLMNOPQRSTUVWXYZ_0123456789=") == std::string::npos &&
                eq != std::string::npos)
                d += "#define " + item.substr(0, eq) + " " + item.substr(eq + 1) + "\n";
            i = j + 1;
        }
    }
    return d;
}

const char *kStubSpmv =
    "extern \"C\" __global__ void __launch_bounds__(SB_WG, SB_SPMV_WAVES(SB_UNROLL)) sb_spmv_rtc(SbSpmvArgs a)\n"
    "{ sb_spmv_body<SB_WG, SB_UNROLL, SB_XCAP>(a); }\n";
const char *kStubSpmmF32 =
    "extern \"C\" __global__ void __launch_bounds__(SB_WG, SB_F32_WAVES(SB_NRHS)) sb_spmm_f32_rtc(SbSpmmArgs a)\n"
    "{ sb_spmm_f32_body<SB_WG, SB_RTMAX, SB_NRHS, SB_XCAP>(a); }\n";
const char *kStubSpmmMfma =
    "extern \"C\" __global__ void __launch_bounds__(SB_WG, SB_MFMA_WAVES) sb_spmm_mfma_rtc(SbSpmmArgs a)\n"
    "{ sb_spmm_mfma_body<SB_WG, SB_NRHS, SB_XCAP, SB_SPLIT>(a); }\n";

int pow2_at_least(int v)
{
    int p = 16;
    while (p < v) p <<= 1;
    return p;
}

}  // namespace

// ---------------------------------------------------------------------------
// plan object

// One launch of a plan or group part: a contiguous range of its tiles and
// the kernel that runs them.  SpMV plans and SpMM plans without MFMA pieces
// have one phase; SpMM plans with MFMA pieces run their MFMA tiles first and
// then the tiles of the other pieces, which add to Y.
struct Phase {
    int32_t t0 = 0, t1 = 0;
    bool rtc = false;
    hipFunction_t fn = nullptr;
    const void *aot = nullptr;
    int grid = 0;
    int32_t spec[6] = {0, 0, 0, 0, 0, 0};   // kind, wg, unroll, xcap, nrhs, rtmax
};

struct SablePlan {
    sable::HostPlan host;
    bool host_only = false;
    bool keep_host = true;
    SablePlanOpts opts{};
    SablePlanStats stats{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // device arrays
    SbTile *d_tiles = nullptr;
    SbPiece *d_pieces = nullptr;
    SbXRange *d_xr = nullptr;
    float *d_dval = nullptr;
    float *d_sval = nullptr;
    uint16_t *d_sidx = nullptr;
    int32_t *d_srowptr = nullptr;
    int32_t *d_xmap = nullptr;
    int32_t *d_queue = nullptr;
    // MFMA tiles split along k: partial-sum scratch and the fixup list
    float *d_part = nullptr;
    SbTile *d_fix_tiles = nullptr;
    int32_t nfix = 0, fix_groups = 1;
    // host staging for SABLE_HOST_PTRS
    float *d_xtmp = nullptr, *d_ytmp = nullptr;
    int64_t xtmp_n = 0, ytmp_n = 0;
    // flush buffer for cold timing
    uint32_t *d_flush = nullptr;
    int64_t flush_n = 0;
    // kernel
    std::vector<Phase> phases;
    int64_t device_bytes = 0;
};

// One launch (per phase) of a group: the members that share every option
// the kernel specialisation depends on (op, RHS width, workgroup size, MFMA
// split, SpMV unroll, hipRTC use).  Their tile and piece descriptors are
// concatenated on the device; their data stays in the members' own arrays,
// reached through a per-matrix buffer table, so a group copies no matrix.
struct GroupPart {
    int32_t op = SABLE_OP_SPMV, n_rhs = 1, wg = 256, mode = SABLE_MODE_QUEUE, grid = 0, ntiles = 0;
    SbTile *d_tiles = nullptr;
    SbPiece *d_pieces = nullptr;
    SbXRange *d_xr = nullptr;
    SbMatBufs *d_mats = nullptr;
    int32_t *d_queue = nullptr;
    SbTile *d_fix_tiles = nullptr;   // the members' split MFMA row tiles (matrix slot in .mat)
    int32_t nfix = 0, fix_groups = 1;
    int32_t ntiles_mfma = 0;         // MFMA tiles (the first phase when nonzero)
    std::vector<Phase> phases;
};

struct SableGroup {
    std::vector<SablePlan *> plans;
    std::vector<GroupPart> parts;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t *d_flush = nullptr;
    int64_t flush_n = 0;
};

namespace {

template <typename T>
int upload(T **dptr, const std::vector<T> &v, int64_t &bytes)
{
    *dptr = nullptr;
    if (v.empty()) return SABLE_OK;
    HIPOK(hipMalloc((void **)dptr, v.size() * sizeof(T)));
    HIPOK(hipMemcpy(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    bytes += (int64_t)(v.size() * sizeof(T));
    return SABLE_OK;
}

#define CHECKRC(x)              \
    do {                        \
        int _rc = (x);          \
        if (_rc) return _rc;    \
    } while (0)

int max_piece_cols(const sable::HostPlan &h)
{
    int m = 1;
    for (auto &p : h.pieces) m = std::max(m, p.ncols);
    return m;
}

// chooses the kernel of one phase (tiles [t0, t1) of the host plan view):
// an RTC specialisation sized for those tiles, or the generic AOT kernel
int select_phase(const sable::HostPlan &h, const SablePlanOpts &o, Phase &ph)
{
    ph.rtc = false;
    ph.fn = nullptr;
    Spec s;
    s.wg = o.wg;
    s.unroll = o.unroll;
    const bool spmm = h.op == SABLE_OP_SPMM;
    bool mfma = false;
    int maxcols = 1, maxstage = 1, rows = 1;
    for (int32_t t = ph.t0; t < ph.t1; ++t) {
        const SbTile &tl = h.tiles[t];
        rows = std::max(rows, tl.nrows);
        for (int32_t k = tl.piece_begin; k < tl.piece_end; ++k) {
            const SbPiece &pc = h.pieces[k];
            maxcols = std::max(maxcols, pc.ncols);
            maxstage = std::max(maxstage, pc.stage_len - 4);
            mfma = mfma || pc.kind == SB_PIECE_MFMA;
        }
    }
    // SpMV: room for the largest staging group the host formed (at most
    // SB_STAGE_CAP + 4 slots, so the buffer never exceeds 4096 + 4)
    // SpMM: the widest piece's staged columns (the MFMA kernel's k-steps
    // read up to 31 past them in its padded xrow; the FP32 kernel's X
    // staging buffer is sized by it, so narrower pieces use less LDS)
    s.xcap = spmm ? std::min(4096, pow2_at_least(mfma ? maxcols + 1 : maxcols))
                  : std::min(sable::SB_STAGE_CAP, std::max(16, pow2_at_least(maxstage)));
    s.split = o.mfma_split == 2 ? 2 : 3;
    if (spmm) {
        s.nrhs = h.n_rhs;
        // FP32 kernel: rows per thread group (the wide shape's are whole row quads)
        const bool wide = SB_F32_WIDE(h.n_rhs);
        const int ngroups = std::max(1, o.wg / (h.n_rhs / SB_F32_COLS(h.n_rhs)));
        const int rt = std::max(1, (rows + ngroups - 1) / ngroups);
        s.rtmax = mfma ? 8 : wide ? (rt + 3) / 4 * 4 : rt;
    }
    if (spmm)
        ph.aot = mfma ? (s.split == 2 ? (const void *)sb_spmm_mfma2_aot : (const void *)sb_spmm_mfma_aot)
                      : (const void *)sb_spmm_f32_aot;
    else ph.aot = (const void *)sb_spmv_aot;
    ph.spec[0] = spmm ? (mfma ? (s.split == 2 ? 3 : 2) : 1) : 0;
    ph.spec[1] = s.wg;
    ph.spec[2] = s.unroll;
    ph.spec[3] = s.xcap;
    ph.spec[4] = s.nrhs;
    ph.spec[5] = s.rtmax;
    if (!o.use_rtc) {
        if (o.wg != 256 || (!spmm && o.unroll != 4)) return fail(SABLE_ERR_UNSUPPORTED, "AOT kernels exist only for wg=256 / unroll=4");
        return SABLE_OK;
    }
    const char *stub = spmm ? (mfma ? kStubSpmmMfma : kStubSpmmF32) : kStubSpmv;
    const char *entry = spmm ? (mfma ? "sb_spmm_mfma_rtc" : "sb_spmm_f32_rtc") : "sb_spmv_rtc";
    int rc = rtc_get(spec_defines(s), stub, entry, &ph.fn);
    if (rc == SABLE_OK) {
        ph.rtc = true;
        return SABLE_OK;
    }
    // fall back to the generic kernel; keep the message for the caller
    if (o.wg != 256 || (!spmm && o.unroll != 4)) return rc;
    fprintf(stderr, "sable: %s; using the generic kernel\n", g_err.c_str());
    return SABLE_OK;
}

// The phases of a plan view: [0, ntiles_mfma) on the MFMA kernel, the rest
// on the FP32 (SpMM) or SpMV kernel.
int select_phases(const sable::HostPlan &h, const SablePlanOpts &o, int32_t ntiles_mfma,
                  std::vector<Phase> &phases)
{
    phases.clear();
    const int32_t n = (int32_t)h.tiles.size();
    if (ntiles_mfma > 0) {
        Phase a;
        a.t0 = 0;
        a.t1 = ntiles_mfma;
        phases.push_back(a);
    }
    if (n > ntiles_mfma || n == 0) {
        Phase b;
        b.t0 = ntiles_mfma;
        b.t1 = n;
        phases.push_back(b);
    }
    for (auto &ph : phases) CHECKRC(select_phase(h, o, ph));
    return SABLE_OK;
}

int occupancy_grid(bool rtc, hipFunction_t fn, const void *aot, int wg, int ntiles, int req)
{
    if (req > 0) return std::max(1, std::min(req, ntiles));
    int per_cu = 0;
    if (rtc) {
        if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, wg, 0) != hipSuccess) per_cu = 0;
    } else {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, aot, wg, 0) != hipSuccess) per_cu = 0;
    }
    if (per_cu <= 0) per_cu = 4;
    int64_t g = (int64_t)per_cu * std::max(g_ncu, 1);
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, ntiles));
}

// Launches one kernel.  With events, e0/e1 take the kernel's own start and
// end times (AMD launch extension), so timings exclude the dispatch latency.
int launch_raw(bool rtc, hipFunction_t fn, const void *aot, int grid, int wg, void *argp,
               hipStream_t st, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    void *args[] = {argp};
    if (rtc) {
        if (e0) {
            HIPOK(hipExtModuleLaunchKernel(fn, (uint32_t)grid * (uint32_t)wg, 1, 1, wg, 1, 1, 0, st, args, nullptr,
                                           e0, e1, 0));
        } else {
            HIPOK(hipModuleLaunchKernel(fn, grid, 1, 1, wg, 1, 1, 0, st, args, nullptr));
        }
    } else if (e0) {
        HIPOK(hipExtLaunchKernel(aot, dim3(grid), dim3(wg), args, 0, st, e0, e1, 0));
    } else {
        HIPOK(hipLaunchKernel(aot, dim3(grid), dim3(wg), args, 0, st));
    }
    return SABLE_OK;
}

// Launches one phase; events (when given) time this launch alone.
int launch_phase(int op, int n_rhs, const Phase &ph, int wg, int mode, const SbTile *tiles,
                 const SbPiece *pieces, const SbXRange *xr, const float *dval, const float *sval,
                 const uint16_t *sidx, const int32_t *srowptr, const int32_t *xmap, int32_t *queue,
                 const SbMatBufs *mats, const float *x, float *y, float *part, hipStream_t st,
                 hipEvent_t e0, hipEvent_t e1)
{
    const int n = ph.t1 - ph.t0;
    if (n <= 0) return SABLE_OK;
    const int grid = mode == SABLE_MODE_GRID ? n : std::min(ph.grid, n);
    if (op == SABLE_OP_SPMV) {
        SbSpmvArgs a{};
        a.tiles = tiles + ph.t0; a.pieces = pieces; a.xr = xr;
        a.dval = dval; a.sval = sval; a.sidx = sidx; a.srowptr = srowptr; a.xmap = xmap;
        a.mats = mats; a.x0 = x; a.y0 = y; a.queue = queue;
        a.ntiles = n; a.mode = mode;
        return launch_raw(ph.rtc, ph.fn, ph.aot, grid, wg, &a, st, e0, e1);
    }
    SbSpmmArgs a{};
    a.tiles = tiles + ph.t0; a.pieces = pieces; a.xr = xr;
    a.dval = dval; a.sval = sval; a.sidx = sidx; a.srowptr = srowptr;
    a.mats = mats; a.x0 = x; a.y0 = y; a.queue = queue; a.part = part;
    a.ntiles = n; a.mode = mode; a.n_rhs = n_rhs;
    return launch_raw(ph.rtc, ph.fn, ph.aot, grid, wg, &a, st, e0, e1);
}

// Launches the fixup of a plan's (or group part's) split MFMA tiles:
// `groups` workgroups per split row tile.
int launch_fixup(const SbTile *fix_tiles, int nfix, int groups, const SbMatBufs *mats, const float *part, float *y,
                 int n_rhs, hipStream_t st)
{
    if (nfix <= 0) return SABLE_OK;
    hipLaunchKernelGGL(sb_mfma_fixup, dim3(nfix, groups), dim3(256), 0, st, fix_tiles, nfix, mats, part, y, n_rhs);
    HIPOK(hipGetLastError());
    return SABLE_OK;
}

// Launches every phase of a plan in order on its stream (the fixup of split
// MFMA tiles right after the MFMA phase), with the partial-sum scratch
// `part`.  With events: a single launch carries them itself (kernel start
// and end); several launches are bracketed by stream events (their gaps
// included).
int plan_launch_with(SablePlan *p, const float *x, float *y, float *part, hipEvent_t e0 = nullptr,
                     hipEvent_t e1 = nullptr)
{
    const size_t np = p->stats.ntiles > 0 ? p->phases.size() : 0;
    const bool attached = np == 1 && p->nfix == 0;
    if (e0 && !attached) HIPOK(hipEventRecord(e0, p->stream));
    for (size_t i = 0; i < np; ++i) {
        const Phase &ph = p->phases[i];
        CHECKRC(launch_phase(p->host.op, p->host.n_rhs, ph, p->opts.wg, p->opts.mode, p->d_tiles, p->d_pieces,
                             p->d_xr, p->d_dval, p->d_sval, p->d_sidx, p->d_srowptr, p->d_xmap, p->d_queue, nullptr,
                             x, y, part, p->stream, attached ? e0 : nullptr, attached ? e1 : nullptr));
        if (i == 0 && p->stats.ntiles_mfma > 0)
            CHECKRC(launch_fixup(p->d_fix_tiles, p->nfix, p->fix_groups, nullptr, part, y, p->host.n_rhs, p->stream));
    }
    if (e1 && !attached) HIPOK(hipEventRecord(e1, p->stream));
    return SABLE_OK;
}

int plan_launch(SablePlan *p, const float *x, float *y, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    return plan_launch_with(p, x, y, p->d_part, e0, e1);
}

int ensure_flush(uint32_t **buf, int64_t *n, int64_t bytes)
{
    const int64_t want = (bytes + 3) / 4;
    if (*n >= want) return SABLE_OK;
    if (*buf) hipFree(*buf);
    *buf = nullptr;
    *n = 0;
    HIPOK(hipMalloc((void **)buf, want * 4));
    *n = want;
    return SABLE_OK;
}

int flush_caches(uint32_t *buf, int64_t n, hipStream_t st, uint32_t salt)
{
    if (!buf || n == 0) return SABLE_OK;
    hipLaunchKernelGGL(sb_fill_u32, dim3(std::max(1, g_ncu) * 8), dim3(256), 0, st, buf, n, salt);
    HIPOK(hipGetLastError());
    return SABLE_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI

extern "C" {

const char *sable_last_error(void) { return g_err.c_str(); }

int sable_init(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(SABLE_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return fail(SABLE_ERR_INVALID, "device index out of range");
    HIPOK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPOK(hipGetDeviceProperties(&prop, device));
    g_device = device;
    g_ncu = prop.multiProcessorCount;
    if (device < 64 && !g_scratch[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return fail(SABLE_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = ptr;
    }
    g_hbm = (int64_t)prop.totalGlobalMem;
    std::string arch = prop.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch = arch.substr(0, colon);
    if (arch.rfind("gfx950", 0) != 0)
        fprintf(stderr, "sable: device is %s; kernels are written for gfx950\n", arch.c_str());
    g_arch = arch;
    return SABLE_OK;
}

int sable_device_info(int *n_cu, int64_t *hbm_bytes, char *arch, int arch_len)
{
    if (g_device < 0) return fail(SABLE_ERR_NO_DEVICE, "sable_init not called");
    if (n_cu) *n_cu = g_ncu;
    if (hbm_bytes) *hbm_bytes = g_hbm;
    if (arch && arch_len > 0) {
        strncpy(arch, g_arch.c_str(), arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    return SABLE_OK;
}

int sable_set_cache_dir(const char *dir)
{
    g_cache_dir = dir ? dir : "";
    return SABLE_OK;
}

// Compiles (and caches) one kernel specialisation without a device: used at
// build time so the code objects ship with the library.  No counterpart in
// the reference (build tooling).  kind: 0 spmv, 1 spmm f32, 2 spmm mfma (split 3),
// 3 spmm mfma (split 2).
int sable_rtc_precompile(int kind, int wg, int unroll, int xcap, int nrhs, int rtmax,
                         const char *arch)
{
    if (arch && *arch) g_arch = arch;
    Spec s;
    s.wg = wg; s.unroll = unroll; s.xcap = xcap; s.nrhs = nrhs; s.rtmax = rtmax;
    s.split = kind == 3 ? 2 : 3;
    const char *stub = kind == 0 ? kStubSpmv : (kind == 1 ? kStubSpmmF32 : kStubSpmmMfma);
    std::vector<char> code;
    std::string path;
    return rtc_code_object(spec_defines(s), stub, code, path);
}

int sable_plan_create(const SableVbr *vbr, const uint8_t *route, const SablePlanOpts *opts_in,
                      SablePlan **out)
{
    if (!vbr || !out) return fail(SABLE_ERR_INVALID, "null argument");
    *out = nullptr;
    SablePlanOpts o = sable::normalize_opts(opts_in);
    std::unique_ptr<SablePlan> p(new SablePlan());
    p->opts = o;
    std::string err;
    int rc;
    try {
        rc = sable::build_host_plan(*vbr, route, o, p->host, err);
    } catch (const std::bad_alloc &) {
        return fail(SABLE_ERR_OOM, "host allocation failed while building the plan");
    }
    if (rc) return fail(rc, err);
    p->stats = p->host.stats;
    p->host_only = o.host_only != 0;
    if (p->host_only) {
        SablePlanOpts no_rtc = o;
        no_rtc.use_rtc = 0;
        no_rtc.wg = 256;       // the spec is recorded, nothing is compiled
        select_phases(p->host, no_rtc, p->stats.ntiles_mfma, p->phases);
        for (auto &ph : p->phases) {
            ph.spec[1] = o.wg;
            ph.spec[2] = o.unroll;
        }
        *out = p.release();
        return SABLE_OK;
    }
    if (g_device < 0) return fail(SABLE_ERR_NO_DEVICE, "sable_init not called");
    CHECKRC(select_phases(p->host, o, p->stats.ntiles_mfma, p->phases));
    const sable::HostPlan &h = p->host;
    int64_t bytes = 0;
    CHECKRC(upload(&p->d_tiles, h.tiles, bytes));
    CHECKRC(upload(&p->d_pieces, h.pieces, bytes));
    CHECKRC(upload(&p->d_xr, h.xr, bytes));
    CHECKRC(upload(&p->d_dval, h.dval, bytes));
    CHECKRC(upload(&p->d_sval, h.sval, bytes));
    CHECKRC(upload(&p->d_sidx, h.sidx, bytes));
    CHECKRC(upload(&p->d_srowptr, h.srowptr, bytes));
    CHECKRC(upload(&p->d_xmap, h.xmap, bytes));
    HIPOK(hipMalloc((void **)&p->d_queue, 2 * sizeof(int32_t)));
    HIPOK(hipMemset(p->d_queue, 0, 2 * sizeof(int32_t)));
    if (!h.fix_tiles.empty()) {
        HIPOK(hipMalloc((void **)&p->d_part, (size_t)h.part_blocks * 1024));
        CHECKRC(upload(&p->d_fix_tiles, h.fix_tiles, bytes));
        p->nfix = (int32_t)h.fix_tiles.size();
        p->fix_groups = sb_fixup_groups(h.fix_tiles.data(), h.fix_tiles.size(), h.n_rhs);
        bytes += h.part_blocks * 1024;
    }
    HIPOK(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    p->own_stream = true;
    p->stats.rtc_used = 1;
    for (auto &ph : p->phases) {
        ph.grid = occupancy_grid(ph.rtc, ph.fn, ph.aot, o.wg, std::max(1, ph.t1 - ph.t0), o.grid);
        p->stats.rtc_used = p->stats.rtc_used && ph.rtc;
    }
    p->stats.grid = p->phases.empty() ? 0 : p->phases[0].grid;
    p->stats.device_bytes = bytes;
    p->device_bytes = bytes;
    // the big arrays live on the device now; keep only descriptors on the host
    p->host.dval = std::vector<float>();
    p->host.sval = std::vector<float>();
    p->host.sidx = std::vector<uint16_t>();
    p->host.srowptr = std::vector<int32_t>();
    p->host.xmap = std::vector<int32_t>();
    p->keep_host = false;
    *out = p.release();
    return SABLE_OK;
}

int sable_plan_destroy(SablePlan *p)
{
    if (!p) return SABLE_OK;
    if (p->stream) hipStreamSynchronize(p->stream);
    void *bufs[] = {p->d_tiles, p->d_pieces, p->d_xr, p->d_dval, p->d_sval, p->d_sidx, p->d_srowptr, p->d_xmap,
                    p->d_queue, p->d_part, p->d_fix_tiles, p->d_xtmp, p->d_ytmp, p->d_flush};
    for (void *b : bufs)
        if (b) hipFree(b);
    if (p->own_stream && p->stream) hipStreamDestroy(p->stream);
    delete p;
    return SABLE_OK;
}

int sable_plan_kernel_spec(const SablePlan *p, int32_t *spec6)
{
    if (!p || !spec6) return fail(SABLE_ERR_INVALID, "null argument");
    if (p->phases.empty()) return fail(SABLE_ERR_INVALID, "plan has no kernel");
    for (int i = 0; i < 6; ++i) spec6[i] = p->phases[0].spec[i];
    return SABLE_OK;
}

int sable_plan_stats(const SablePlan *p, SablePlanStats *out)
{
    if (!p || !out) return fail(SABLE_ERR_INVALID, "null argument");
    *out = p->stats;
    return SABLE_OK;
}

int sable_plan_set_stream(SablePlan *p, void *stream)
{
    if (!p || p->host_only) return fail(SABLE_ERR_INVALID, "invalid plan");
    HIPOK(hipStreamSynchronize(p->stream));
    if (p->own_stream) hipStreamDestroy(p->stream);
    p->stream = (hipStream_t)stream;
    p->own_stream = false;
    return SABLE_OK;
}

int sable_plan_export(const SablePlan *p, int which, void *dst, int64_t *bytes)
{
    if (!p || !bytes) return fail(SABLE_ERR_INVALID, "null argument");
    const sable::HostPlan &h = p->host;
    const void *hsrc = nullptr;
    const void *dsrc = nullptr;
    int64_t n = 0;
    switch (which) {
    case 0: hsrc = h.tiles.data(); n = h.tiles.size() * sizeof(SbTile); break;
    case 1: hsrc = h.pieces.data(); n = h.pieces.size() * sizeof(SbPiece); break;
    case 2: hsrc = h.xr.data(); n = h.xr.size() * sizeof(SbXRange); break;
    case 3: hsrc = h.dval.data(); dsrc = p->d_dval; n = (p->stats.dense_padded + p->stats.mfma_padded) * 4; break;
    case 4: hsrc = h.sval.data(); dsrc = p->d_sval; n = p->stats.sparse_padded * 4; break;
    case 5: hsrc = h.sidx.data(); dsrc = p->d_sidx; n = p->stats.sparse_padded * 2; break;
    case 6: hsrc = h.srowptr.data(); dsrc = p->d_srowptr; n = p->stats.rowptr_entries * 4; break;
    case 7: hsrc = h.xmap.data(); dsrc = p->d_xmap; n = p->stats.xmap_entries * 4; break;
    default: return fail(SABLE_ERR_INVALID, "unknown export id");
    }
    *bytes = n;
    if (!dst || n == 0) return SABLE_OK;
    if (p->keep_host || which <= 2) {
        std::memcpy(dst, hsrc, n);
    } else {
        HIPOK(hipMemcpy(dst, dsrc, n, hipMemcpyDeviceToHost));
    }
    return SABLE_OK;
}

static int run_op(SablePlan *p, const float *x, float *y, int flags, int64_t nx, int64_t ny)
{
    if (!p) return fail(SABLE_ERR_INVALID, "null plan");
    if (p->host_only) return fail(SABLE_ERR_INVALID, "host-only plan cannot run");
    if (!x || !y) return fail(SABLE_ERR_INVALID, "null x/y");
    const float *dx = x;
    float *dy = y;
    if (flags & SABLE_HOST_PTRS) {
        if (p->xtmp_n < nx) {
            if (p->d_xtmp) hipFree(p->d_xtmp);
            p->d_xtmp = nullptr;
            HIPOK(hipMalloc((void **)&p->d_xtmp, std::max<int64_t>(nx, 1) * 4));
            p->xtmp_n = nx;
        }
        if (p->ytmp_n < ny) {
            if (p->d_ytmp) hipFree(p->d_ytmp);
            p->d_ytmp = nullptr;
            HIPOK(hipMalloc((void **)&p->d_ytmp, std::max<int64_t>(ny, 1) * 4));
            p->ytmp_n = ny;
        }
        HIPOK(hipMemcpyAsync(p->d_xtmp, x, nx * 4, hipMemcpyHostToDevice, p->stream));
        dx = p->d_xtmp;
        dy = p->d_ytmp;
    }
    CHECKRC(plan_launch(p, dx, dy));
    HIPOK(hipGetLastError());
    if (flags & SABLE_HOST_PTRS) {
        HIPOK(hipMemcpyAsync(y, dy, ny * 4, hipMemcpyDeviceToHost, p->stream));
        HIPOK(hipStreamSynchronize(p->stream));
    } else if (!(flags & SABLE_NO_SYNC)) {
        HIPOK(hipStreamSynchronize(p->stream));
    }
    return SABLE_OK;
}

int sable_spmv(SablePlan *p, const float *x, float *y, int flags)
{
    if (p && p->host.op != SABLE_OP_SPMV) return fail(SABLE_ERR_INVALID, "plan was built for SpMM");
    return run_op(p, x, y, flags, p ? p->stats.ncols : 0, p ? p->stats.nrows : 0);
}

int sable_spmm(SablePlan *p, const float *X, float *Y, int flags)
{
    if (p && p->host.op != SABLE_OP_SPMM) return fail(SABLE_ERR_INVALID, "plan was built for SpMV");
    return run_op(p, X, Y, flags, p ? (int64_t)p->stats.ncols * p->stats.n_rhs : 0,
                  p ? (int64_t)p->stats.nrows * p->stats.n_rhs : 0);
}

// Y = A X like sable_spmm with device X/Y, allocating nothing: the split MFMA
// tiles' partial sums live in the file-scope `scratch` array (through its
// per-device address) instead of the plan's own buffer.  Nothing in it needs
// to start at zero: every partial is written before the fixup reads it in
// the same call.  Calls that overlap on one GPU share the array and race.
int sable_spmm_noalloc(SablePlan *p, const float *X, float *Y, int flags)
{
    if (!p || p->host_only) return fail(SABLE_ERR_INVALID, "invalid plan");
    if (p->host.op != SABLE_OP_SPMM) return fail(SABLE_ERR_INVALID, "plan was built for SpMV");
    if (!X || !Y) return fail(SABLE_ERR_INVALID, "null X/Y");
    if (flags & SABLE_HOST_PTRS) return fail(SABLE_ERR_INVALID, "sable_spmm_noalloc takes device pointers only");
    if (g_device < 0 || g_device >= 64 || !g_scratch[g_device]) return fail(SABLE_ERR_NO_DEVICE, "sable_init not called");
    const int64_t need = p->host.part_blocks * 1024;
    if (need > SABLE_NOALLOC_SCRATCH_BYTES)
        return fail(SABLE_ERR_UNSUPPORTED, "plan needs " + std::to_string(need) + " bytes of scratch; "
                    "sable_spmm_noalloc holds at most " + std::to_string(SABLE_NOALLOC_SCRATCH_BYTES));
    CHECKRC(plan_launch_with(p, X, Y, (float *)g_scratch[g_device]));
    HIPOK(hipGetLastError());
    if (!(flags & SABLE_NO_SYNC)) HIPOK(hipStreamSynchronize(p->stream));
    return SABLE_OK;
}

int sable_time(SablePlan *p, const float *x, float *y, int warmup, int reps, int64_t flush_bytes,
               float *times_ms)
{
    if (!p || p->host_only) return fail(SABLE_ERR_INVALID, "invalid plan");
    if (reps < 0 || warmup < 0 || (reps > 0 && !times_ms)) return fail(SABLE_ERR_INVALID, "bad reps");
    if (flush_bytes > 0) CHECKRC(ensure_flush(&p->d_flush, &p->flush_n, flush_bytes));
    for (int i = 0; i < warmup; ++i) {
        if (flush_bytes > 0) CHECKRC(flush_caches(p->d_flush, p->flush_n, p->stream, (uint32_t)i));
        CHECKRC(plan_launch(p, x, y));
    }
    hipEvent_t e0, e1;
    HIPOK(hipEventCreate(&e0));
    HIPOK(hipEventCreate(&e1));
    int rc = SABLE_OK;
    for (int r = 0; r < reps && rc == SABLE_OK; ++r) {
        if (flush_bytes > 0) rc = flush_caches(p->d_flush, p->flush_n, p->stream, (uint32_t)(r + 7));
        if (rc) break;
        rc = plan_launch(p, x, y, e0, e1);   // events on the kernel itself
        if (rc) break;
        if (hipEventSynchronize(e1) != hipSuccess) { rc = fail(SABLE_ERR_HIP, "event sync"); break; }
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        times_ms[r] = ms;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    HIPOK(hipStreamSynchronize(p->stream));
    return SABLE_OK;
}

// ---- groups: many plans in persistent launches, captured in one hipGraph

namespace {

void free_part(GroupPart &q)
{
    void *bufs[] = {q.d_tiles, q.d_pieces, q.d_xr, q.d_mats, q.d_queue, q.d_fix_tiles};
    for (void *b : bufs)
        if (b) hipFree(b);
    q = GroupPart();
}

// Builds one part from members idx (all sharing the kernel options): their
// tile and piece descriptors concatenated in one global LPT order, with the
// tiles' matrix slots pointing into a table of the members' own device
// arrays and x/y, one kernel specialisation.
int build_part(SablePlan *const *plans, const std::vector<int> &idx, const float *const *x, float *const *y,
               int mode, int grid, GroupPart &q)
{
    SablePlan *p0 = plans[idx[0]];
    q.op = p0->host.op;
    q.n_rhs = p0->host.n_rhs;
    q.wg = p0->opts.wg;
    q.mode = (mode < 0 || mode > 2) ? SABLE_MODE_QUEUE : mode;
    std::vector<SbTile> tiles;
    std::vector<SbPiece> pieces;
    std::vector<SbXRange> xr;
    std::vector<double> cost;
    std::vector<int32_t> phase;
    std::vector<SbMatBufs> mats;
    std::vector<SbTile> fixes;
    for (size_t m = 0; m < idx.size(); ++m) {
        SablePlan *p = plans[idx[m]];
        if (pieces.size() + p->host.pieces.size() > (size_t)INT32_MAX ||
            xr.size() + p->host.xr.size() > (size_t)INT32_MAX)
            return fail(SABLE_ERR_UNSUPPORTED, "group descriptor arrays exceed 2^31 entries");
        const int32_t pbase = (int32_t)pieces.size(), xbase = (int32_t)xr.size();
        for (auto pc : p->host.pieces) {
            pc.xr_begin += xbase;
            pc.xr_end += xbase;
            if (pc.stage_len > 0) pc.grp_end += pbase;   // first piece of a staging group
            pieces.push_back(pc);                        // data, row pointer and map offsets stay the member's
        }
        xr.insert(xr.end(), p->host.xr.begin(), p->host.xr.end());
        for (size_t t = 0; t < p->host.tiles.size(); ++t) {
            SbTile tt = p->host.tiles[t];
            tt.piece_begin += pbase;
            tt.piece_end += pbase;
            tt.mat = (int32_t)m;
            tiles.push_back(tt);
            cost.push_back(p->host.tile_cost[t]);
            phase.push_back((int32_t)t < p->stats.ntiles_mfma ? 0 : 1);
        }
        SbMatBufs b{};
        b.dval = p->d_dval;
        b.sval = p->d_sval;
        b.sidx = p->d_sidx;
        b.srowptr = p->d_srowptr;
        b.xmap = p->d_xmap;
        b.x = x[idx[m]];
        b.y = y[idx[m]];
        b.part = p->d_part;
        mats.push_back(b);
        for (SbTile ft : p->host.fix_tiles) {
            ft.mat = (int32_t)m;
            fixes.push_back(ft);
        }
    }
    // global LPT order: MFMA tiles first (first launch), largest first in each phase
    std::vector<int32_t> order(tiles.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int32_t)i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        if (phase[a] != phase[b]) return phase[a] < phase[b];
        return cost[a] > cost[b];
    });
    int32_t nm = 0;
    for (int32_t ph : phase) nm += (ph == 0);
    q.ntiles_mfma = nm;
    std::vector<SbTile> sorted(tiles.size());
    for (size_t i = 0; i < order.size(); ++i) sorted[i] = tiles[order[i]];
    q.ntiles = (int32_t)sorted.size();
    int64_t bytes = 0;
    CHECKRC(upload(&q.d_tiles, sorted, bytes));
    CHECKRC(upload(&q.d_pieces, pieces, bytes));
    CHECKRC(upload(&q.d_xr, xr, bytes));
    CHECKRC(upload(&q.d_mats, mats, bytes));
    CHECKRC(upload(&q.d_fix_tiles, fixes, bytes));
    q.nfix = (int32_t)fixes.size();
    q.fix_groups = sb_fixup_groups(fixes.data(), fixes.size(), q.n_rhs);
    HIPOK(hipMalloc((void **)&q.d_queue, 2 * sizeof(int32_t)));
    HIPOK(hipMemset(q.d_queue, 0, 2 * sizeof(int32_t)));
    // kernel: the specialisation machinery over a host plan view of the part
    sable::HostPlan view;
    view.op = q.op;
    view.n_rhs = q.n_rhs;
    view.tiles = sorted;
    view.pieces = pieces;
    CHECKRC(select_phases(view, p0->opts, nm, q.phases));
    for (auto &ph : q.phases)
        ph.grid = occupancy_grid(ph.rtc, ph.fn, ph.aot, q.wg, std::max(1, ph.t1 - ph.t0), grid);
    return SABLE_OK;
}

int launch_part(const GroupPart &q, hipStream_t st)
{
    const bool has_mfma_phase = q.ntiles_mfma > 0;
    for (size_t i = 0; i < q.phases.size(); ++i) {
        CHECKRC(launch_phase(q.op, q.n_rhs, q.phases[i], q.wg, q.mode, q.d_tiles, q.d_pieces, q.d_xr, nullptr, nullptr,
                             nullptr, nullptr, nullptr, q.d_queue, q.d_mats, nullptr, nullptr, nullptr, st, nullptr,
                             nullptr));
        if (i == 0 && has_mfma_phase) CHECKRC(launch_fixup(q.d_fix_tiles, q.nfix, q.fix_groups, q.d_mats, nullptr, nullptr, q.n_rhs, st));
    }
    return SABLE_OK;
}

}  // namespace

int sable_group_create(SablePlan *const *plans, int nplans, const float *const *x, float *const *y,
                       int mode, int grid, SableGroup **out)
{
    if (!plans || nplans <= 0 || !x || !y || !out) return fail(SABLE_ERR_INVALID, "null argument");
    *out = nullptr;
    if (g_device < 0) return fail(SABLE_ERR_NO_DEVICE, "sable_init not called");
    for (int i = 0; i < nplans; ++i)
        if (!plans[i] || plans[i]->host_only) return fail(SABLE_ERR_INVALID, "group member must be a device plan");
    // a plan's split MFMA tiles own one set of partial sums, so a plan
    // listed twice would have its two copies' chunks share them
    for (int i = 0; i < nplans; ++i)
        for (int j = 0; j < i; ++j)
            if (plans[i] == plans[j] && plans[i]->d_part)
                return fail(SABLE_ERR_INVALID, "a plan with split MFMA tiles is listed twice in the group");
    std::unique_ptr<SableGroup> g(new SableGroup());
    // members sharing every option of the kernel specialisation form one
    // part, in order of first appearance (a part runs its first member's
    // specialisation, so a member never runs another split or unroll)
    std::vector<std::vector<int>> members;
    std::vector<std::array<int, 6>> keys;
    for (int i = 0; i < nplans; ++i) {
        const SablePlanOpts &o = plans[i]->opts;
        const bool spmm = plans[i]->host.op == SABLE_OP_SPMM;
        const std::array<int, 6> k = {plans[i]->host.op, plans[i]->host.n_rhs, o.wg, spmm ? o.mfma_split : 0,
                                      spmm ? 0 : o.unroll, o.use_rtc};
        size_t j = 0;
        while (j < keys.size() && keys[j] != k) ++j;
        if (j == keys.size()) {
            keys.push_back(k);
            members.emplace_back();
        }
        members[j].push_back(i);
        g->plans.push_back(plans[i]);
    }
    g->parts.resize(members.size());
    for (size_t j = 0; j < members.size(); ++j) {
        const int rc = build_part(plans, members[j], x, y, mode, grid, g->parts[j]);
        if (rc) {
            for (auto &q : g->parts) free_part(q);
            return rc;
        }
    }
    HIPOK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    // capture every part's persistent launch, one after the other
    HIPOK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    int rc = SABLE_OK;
    for (auto &q : g->parts) {
        rc = launch_part(q, g->stream);
        if (rc) break;
    }
    hipGraph_t graph = nullptr;
    hipError_t ce = hipStreamEndCapture(g->stream, &graph);
    if (rc == SABLE_OK && ce != hipSuccess)
        rc = fail(SABLE_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    if (rc == SABLE_OK) {
        g->graph = graph;
        if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
            rc = fail(SABLE_ERR_HIP, "hipGraphInstantiate failed");
    } else if (graph) {
        hipGraphDestroy(graph);
    }
    if (rc) {
        sable_group_destroy(g.release());
        return rc;
    }
    *out = g.release();
    return SABLE_OK;
}

int sable_group_launch(SableGroup *g, int sync)
{
    if (!g || !g->exec) return fail(SABLE_ERR_INVALID, "invalid group");
    HIPOK(hipGraphLaunch(g->exec, g->stream));
    if (sync) HIPOK(hipStreamSynchronize(g->stream));
    return SABLE_OK;
}

int sable_group_time(SableGroup *g, int warmup, int reps, int64_t flush_bytes, float *times_ms)
{
    if (!g || !g->exec) return fail(SABLE_ERR_INVALID, "invalid group");
    if (reps > 0 && !times_ms) return fail(SABLE_ERR_INVALID, "null times");
    if (flush_bytes > 0) CHECKRC(ensure_flush(&g->d_flush, &g->flush_n, flush_bytes));
    for (int i = 0; i < warmup; ++i) {
        if (flush_bytes > 0) CHECKRC(flush_caches(g->d_flush, g->flush_n, g->stream, (uint32_t)i));
        HIPOK(hipGraphLaunch(g->exec, g->stream));
    }
    hipEvent_t e0, e1;
    HIPOK(hipEventCreate(&e0));
    HIPOK(hipEventCreate(&e1));
    for (int r = 0; r < reps; ++r) {
        if (flush_bytes > 0) CHECKRC(flush_caches(g->d_flush, g->flush_n, g->stream, (uint32_t)(r + 3)));
        HIPOK(hipEventRecord(e0, g->stream));
        HIPOK(hipGraphLaunch(g->exec, g->stream));
        HIPOK(hipEventRecord(e1, g->stream));
        HIPOK(hipEventSynchronize(e1));
        hipEventElapsedTime(&times_ms[r], e0, e1);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return SABLE_OK;
}

int sable_time_plans(SablePlan *const *plans, int nplans, const float *const *x, float *const *y, int warmup,
                     int reps, int64_t flush_bytes, float *times_ms)
{
    if (!plans || nplans <= 0 || !x || !y) return fail(SABLE_ERR_INVALID, "null argument");
    if (reps < 0 || warmup < 0 || (reps > 0 && !times_ms)) return fail(SABLE_ERR_INVALID, "bad reps");
    for (int i = 0; i < nplans; ++i)
        if (!plans[i] || plans[i]->host_only) return fail(SABLE_ERR_INVALID, "every plan must be a device plan");
    SablePlan *p0 = plans[0];
    hipStream_t st = p0->stream;
    // every launch goes to the first plan's stream, back to back
    auto run_all = [&]() -> int {
        for (int i = 0; i < nplans; ++i) {
            hipStream_t keep = plans[i]->stream;
            plans[i]->stream = st;
            const int rc = plan_launch(plans[i], x[i], y[i]);
            plans[i]->stream = keep;
            if (rc) return rc;
        }
        return SABLE_OK;
    };
    if (flush_bytes > 0) CHECKRC(ensure_flush(&p0->d_flush, &p0->flush_n, flush_bytes));
    for (int i = 0; i < warmup; ++i) {
        if (flush_bytes > 0) CHECKRC(flush_caches(p0->d_flush, p0->flush_n, st, (uint32_t)i));
        CHECKRC(run_all());
    }
    hipEvent_t e0, e1;
    HIPOK(hipEventCreate(&e0));
    HIPOK(hipEventCreate(&e1));
    int rc = SABLE_OK;
    for (int r = 0; r < reps && rc == SABLE_OK; ++r) {
        if (flush_bytes > 0) rc = flush_caches(p0->d_flush, p0->flush_n, st, (uint32_t)(r + 11));
        if (rc) break;
        hipEventRecord(e0, st);
        rc = run_all();
        if (rc) break;
        hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess) { rc = fail(SABLE_ERR_HIP, "event sync"); break; }
        hipEventElapsedTime(&times_ms[r], e0, e1);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    HIPOK(hipStreamSynchronize(st));
    return SABLE_OK;
}

int sable_group_destroy(SableGroup *g)
{
    if (!g) return SABLE_OK;
    if (g->stream) hipStreamSynchronize(g->stream);
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    for (auto &q : g->parts) free_part(q);
    if (g->d_flush) hipFree(g->d_flush);
    if (g->stream) hipStreamDestroy(g->stream);
    delete g;
    return SABLE_OK;
}

// ---- router MLP

int sable_mlp_forward(const float *params, int F, int H, int K, const float *X, int n, float *out,
                      float *hidden)
{
    if (g_device < 0) return fail(SABLE_ERR_NO_DEVICE, "sable_init not called");
    if (F <= 0 || H <= 0 || K <= 0 || n < 0 || H > 256 || F > 64 || K > 64)
        return fail(SABLE_ERR_INVALID, "MLP sizes out of range (F<=64, H<=256, K<=64)");
    if (n == 0) return SABLE_OK;
    hipLaunchKernelGGL(sb_mlp_forward_kernel, dim3((n + 255) / 256), dim3(256), 0, nullptr, params, F, H, K,
                       X, n, out, hidden);
    HIPOK(hipGetLastError());
    HIPOK(hipDeviceSynchronize());
    return SABLE_OK;
}

int sable_mlp_backward(const float *params, int F, int H, int K, const float *X, int n,
                       const float *hidden, const float *dout, float *grad)
{
    if (g_device < 0) return fail(SABLE_ERR_NO_DEVICE, "sable_init not called");
    if (F <= 0 || H <= 0 || K <= 0 || n < 0 || H > 256 || F > 64 || K > 64)
        return fail(SABLE_ERR_INVALID, "MLP sizes out of range (F<=64, H<=256, K<=64)");
    const int64_t np = (int64_t)H * F + H + (int64_t)K * H + K;
    // one workgroup per parameter chunk; deterministic fixed-order sums over samples
    hipLaunchKernelGGL(sb_mlp_backward_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, nullptr, params,
                       F, H, K, X, n, hidden, dout, grad);
    HIPOK(hipGetLastError());
    HIPOK(hipDeviceSynchronize());
    return SABLE_OK;
}

int sable_mlp_mse_grad(const float *out, const float *target, const float *mask, int64_t n, float scale,
                       float *dout)
{
    if (g_device < 0) return fail(SABLE_ERR_NO_DEVICE, "sable_init not called");
    if (n <= 0) return SABLE_OK;
    hipLaunchKernelGGL(sb_mlp_mse_grad_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, out,
                       target, mask, n, scale, dout);
    HIPOK(hipGetLastError());
    HIPOK(hipDeviceSynchronize());
    return SABLE_OK;
}

int sable_adam_step(float *params, const float *grad, float *m, float *v, int64_t n, float lr, float beta1,
                    float beta2, float eps, int step)
{
    if (g_device < 0) return fail(SABLE_ERR_NO_DEVICE, "sable_init not called");
    if (n <= 0 || step <= 0) return fail(SABLE_ERR_INVALID, "bad Adam arguments");
    hipLaunchKernelGGL(sb_adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, params, grad, m, v,
                       n, lr, beta1, beta2, eps, step);
    HIPOK(hipGetLastError());
    HIPOK(hipDeviceSynchronize());
    return SABLE_OK;
}

// ---- bandwidth probe

int sable_bandwidth_probe(int64_t bytes, int reps, float *read_gbs, float *copy_gbs)
{
    if (g_device < 0) return fail(SABLE_ERR_NO_DEVICE, "sable_init not called");
    if (bytes < (1 << 20) || reps <= 0) return fail(SABLE_ERR_INVALID, "bytes >= 1 MiB and reps > 0 needed");
    const int64_t n4 = bytes / 16;
    // 512 KiB per workgroup, at least 8 workgroups per CU
    int64_t chunk4 = (512 << 10) / 16;
    const int64_t min_grid = 8LL * std::max(g_ncu, 1);
    if ((n4 + chunk4 - 1) / chunk4 < min_grid) chunk4 = std::max<int64_t>(256, (n4 + min_grid - 1) / min_grid);
    const int grid = (int)((n4 + chunk4 - 1) / chunk4);
    // every timed run starts cold, like the library's timings: 1 GiB is
    // overwritten first, so the caches hold none of the probe's buffers
    uint32_t *fl = nullptr;
    int64_t fl_n = 0;
    CHECKRC(ensure_flush(&fl, &fl_n, 1LL << 30));
    sb_f4 *a = nullptr, *b = nullptr;
    float *out = nullptr;
    HIPOK(hipMalloc((void **)&a, n4 * 16));
    HIPOK(hipMalloc((void **)&b, n4 * 16));
    HIPOK(hipMalloc((void **)&out, (size_t)grid * 256 * 4));
    HIPOK(hipMemset(a, 0, n4 * 16));
    hipEvent_t e0, e1;
    HIPOK(hipEventCreate(&e0));
    HIPOK(hipEventCreate(&e1));
    float best_r = 0.f, best_c = 0.f;
    for (int r = 0; r < reps + 1; ++r) {
        float ms = 0.f;
        flush_caches(fl, fl_n, nullptr, (uint32_t)r);
        hipEventRecord(e0, nullptr);
        hipLaunchKernelGGL(sb_read_probe, dim3(grid), dim3(256), 0, nullptr, a, n4, chunk4, out);
        hipEventRecord(e1, nullptr);
        HIPOK(hipEventSynchronize(e1));
        hipEventElapsedTime(&ms, e0, e1);
        if (r) best_r = std::max(best_r, (float)(n4 * 16 / (ms * 1e-3) / 1e9));
        flush_caches(fl, fl_n, nullptr, (uint32_t)(r + 1000));
        hipEventRecord(e0, nullptr);
        hipLaunchKernelGGL(sb_copy_probe, dim3(grid), dim3(256), 0, nullptr, a, b, n4, chunk4);
        hipEventRecord(e1, nullptr);
        HIPOK(hipEventSynchronize(e1));
        hipEventElapsedTime(&ms, e0, e1);
        if (r) best_c = std::max(best_c, (float)(2 * n4 * 16 / (ms * 1e-3) / 1e9));
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    hipFree(out);
    hipFree(fl);
    if (read_gbs) *read_gbs = best_r;
    if (copy_gbs) *copy_gbs = best_c;
    return SABLE_OK;
}

// ---- block partitioner: adjacent-line similarity and cut flags

namespace {

struct DevBufs {
    std::vector<void *> p;
    ~DevBufs()
    {
        for (void *q : p) hipFree(q);
    }
    int alloc(void **out, int64_t bytes)
    {
        void *q = nullptr;
        HIPOK(hipMalloc(&q, std::max<int64_t>(bytes, 1)));
        p.push_back(q);
        *out = q;
        return SABLE_OK;
    }
};

}  // namespace

int sable_axis_cuts(const int64_t *indptr, const int32_t *indices, const double *data, int64_t nlines,
                    int64_t line_len, int kind, double threshold, uint8_t *cut, double *sim, float *kernel_ms)
{
    if (g_device < 0) return fail(SABLE_ERR_NO_DEVICE, "sable_init not called");
    if (nlines < 0 || line_len < 0 || !indptr || !cut) return fail(SABLE_ERR_INVALID, "bad axis_cuts arguments");
    if (kind < SB_SIM_S2 || kind > SB_SIM_S1) return fail(SABLE_ERR_INVALID, "unknown similarity kind");
    // the kernels read only inside [indptr[i], indptr[i+1]): check the bounds
    if (indptr[0] != 0) return fail(SABLE_ERR_INVALID, "indptr[0] must be 0");
    for (int64_t i = 0; i < nlines; ++i)
        if (indptr[i + 1] < indptr[i]) return fail(SABLE_ERR_INVALID, "indptr must be nondecreasing");
    const int64_t nnz = indptr[nlines];
    if (nnz > 0 && (!indices || !data)) return fail(SABLE_ERR_INVALID, "null indices/data");
    if (kernel_ms) *kernel_ms = 0.f;
    if (nlines == 0) {
        cut[0] = 1;
        return SABLE_OK;
    }
    DevBufs b;
    int64_t *d_ip = nullptr;
    int32_t *d_ix = nullptr;
    double *d_v = nullptr, *d_sim = nullptr;
    uint8_t *d_cut = nullptr;
    CHECKRC(b.alloc((void **)&d_ip, (nlines + 1) * sizeof(int64_t)));
    CHECKRC(b.alloc((void **)&d_ix, nnz * sizeof(int32_t)));
    CHECKRC(b.alloc((void **)&d_v, nnz * sizeof(double)));
    CHECKRC(b.alloc((void **)&d_sim, nlines * sizeof(double)));
    CHECKRC(b.alloc((void **)&d_cut, nlines + 1));
    HIPOK(hipMemcpy(d_ip, indptr, (nlines + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    if (nnz > 0) {
        HIPOK(hipMemcpy(d_ix, indices, nnz * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPOK(hipMemcpy(d_v, data, nnz * sizeof(double), hipMemcpyHostToDevice));
    }
    hipEvent_t e0, e1;
    HIPOK(hipEventCreate(&e0));
    HIPOK(hipEventCreate(&e1));
    hipEventRecord(e0, nullptr);
    if (nlines > 1)
        hipLaunchKernelGGL(sb_adjacent_sim, dim3((unsigned)((nlines - 1 + 3) / 4)), dim3(256), 0, nullptr, d_ip,
                           d_ix, d_v, nlines, line_len, kind, d_sim);
    hipLaunchKernelGGL(sb_cut_flags, dim3((unsigned)((nlines + 1 + 255) / 256)), dim3(256), 0, nullptr, d_ip, d_sim,
                       nlines, threshold, d_cut);
    hipEventRecord(e1, nullptr);
    const hipError_t le = hipGetLastError();
    const hipError_t se = hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    HIPOK(le);
    HIPOK(se);
    if (kernel_ms) *kernel_ms = ms;
    HIPOK(hipMemcpy(cut, d_cut, nlines + 1, hipMemcpyDeviceToHost));
    if (sim && nlines > 1) HIPOK(hipMemcpy(sim, d_sim, (nlines - 1) * sizeof(double), hipMemcpyDeviceToHost));
    return SABLE_OK;
}

// ---- block materialisation (inspector)

struct SableInspect {
    int64_t nblocks = 0, nval = 0;
    std::vector<int64_t> bpntrb, bpntre;
    std::vector<int64_t> rpntr, cpntr;   // the partitions it was built for
    double *d_val = nullptr;
    int64_t *d_indx = nullptr, *d_bindx = nullptr;
    int32_t *d_bnnz = nullptr;
    ~SableInspect()
    {
        hipFree(d_val);
        hipFree(d_indx);
        hipFree(d_bindx);
        hipFree(d_bnnz);
    }
};

namespace {

int check_partition(const int64_t *part, int32_t nparts, int64_t n, const char *name)
{
    if (!part || nparts < 1) return fail(SABLE_ERR_INVALID, std::string(name) + ": need at least one part");
    if (part[0] != 0 || part[nparts] != n)
        return fail(SABLE_ERR_INVALID, std::string(name) + " must start at 0 and end at the matrix size");
    for (int32_t p = 0; p < nparts; ++p)
        if (part[p + 1] <= part[p]) return fail(SABLE_ERR_INVALID, std::string(name) + " must strictly increase");
    return SABLE_OK;
}

int scan_exclusive(int64_t *d_in, int64_t n, int64_t *d_out, DevBufs &b)
{
    const int64_t nblk = (n + SB_SCAN_BLOCK - 1) / SB_SCAN_BLOCK;
    int64_t *d_tot = nullptr;
    CHECKRC(b.alloc((void **)&d_tot, nblk * sizeof(int64_t)));
    hipLaunchKernelGGL(sb_scan_blocks, dim3((unsigned)nblk), dim3(SB_SCAN_BLOCK), 0, nullptr, d_in, n, d_out, d_tot);
    hipLaunchKernelGGL(sb_scan_totals, dim3(1), dim3(SB_SCAN_BLOCK), 0, nullptr, d_tot, nblk);
    hipLaunchKernelGGL(sb_scan_add, dim3((unsigned)nblk), dim3(SB_SCAN_BLOCK), 0, nullptr, d_out, n, d_tot);
    HIPOK(hipGetLastError());
    return SABLE_OK;
}

}  // namespace

int sable_inspect_create(const int64_t *indptr, const int32_t *indices, const double *data, int64_t nrows,
                         int64_t ncols, const int64_t *rpntr, int32_t nbr, const int64_t *cpntr, int32_t nbc,
                         SableInspect **out, int64_t *nblocks, int64_t *nval, float *device_ms)
{
    if (g_device < 0) return fail(SABLE_ERR_NO_DEVICE, "sable_init not called");
    if (!out || !indptr || nrows < 1 || ncols < 1) return fail(SABLE_ERR_INVALID, "bad inspect arguments");
    *out = nullptr;
    CHECKRC(check_partition(rpntr, nbr, nrows, "rpntr"));
    CHECKRC(check_partition(cpntr, nbc, ncols, "cpntr"));
    // every kernel index is bounded by these checks
    if (indptr[0] != 0) return fail(SABLE_ERR_INVALID, "indptr[0] must be 0");
    for (int64_t j = 0; j < ncols; ++j)
        if (indptr[j + 1] < indptr[j]) return fail(SABLE_ERR_INVALID, "indptr must be nondecreasing");
    const int64_t nnz = indptr[ncols];
    if (nnz > 0 && (!indices || !data)) return fail(SABLE_ERR_INVALID, "null indices/data");
    for (int64_t k = 0; k < nnz; ++k)
        if (indices[k] < 0 || indices[k] >= nrows) return fail(SABLE_ERR_INVALID, "row index out of range");
    std::unique_ptr<SableInspect> h(new SableInspect());
    h->rpntr.assign(rpntr, rpntr + nbr + 1);
    h->cpntr.assign(cpntr, cpntr + nbc + 1);
    DevBufs b;
    const int64_t rw = (nbc + 31) / 32;
    int64_t *d_ip = nullptr, *d_rp = nullptr, *d_cp = nullptr, *d_rowbeg = nullptr, *d_size = nullptr;
    int32_t *d_ix = nullptr, *d_col = nullptr, *d_rmap = nullptr, *d_cmap = nullptr, *d_wpre = nullptr,
            *d_rowcnt = nullptr;
    double *d_v = nullptr;
    uint32_t *d_bits = nullptr;
    CHECKRC(b.alloc((void **)&d_ip, (ncols + 1) * sizeof(int64_t)));
    CHECKRC(b.alloc((void **)&d_ix, nnz * sizeof(int32_t)));
    CHECKRC(b.alloc((void **)&d_v, nnz * sizeof(double)));
    CHECKRC(b.alloc((void **)&d_col, nnz * sizeof(int32_t)));
    CHECKRC(b.alloc((void **)&d_rp, (nbr + 1) * sizeof(int64_t)));
    CHECKRC(b.alloc((void **)&d_cp, (nbc + 1) * sizeof(int64_t)));
    CHECKRC(b.alloc((void **)&d_rmap, nrows * sizeof(int32_t)));
    CHECKRC(b.alloc((void **)&d_cmap, ncols * sizeof(int32_t)));
    CHECKRC(b.alloc((void **)&d_bits, (int64_t)nbr * rw * sizeof(uint32_t)));
    CHECKRC(b.alloc((void **)&d_wpre, (int64_t)nbr * rw * sizeof(int32_t)));
    CHECKRC(b.alloc((void **)&d_rowcnt, nbr * sizeof(int32_t)));
    CHECKRC(b.alloc((void **)&d_rowbeg, nbr * sizeof(int64_t)));
    HIPOK(hipMemcpy(d_ip, indptr, (ncols + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    if (nnz > 0) {
        HIPOK(hipMemcpy(d_ix, indices, nnz * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPOK(hipMemcpy(d_v, data, nnz * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPOK(hipMemcpy(d_rp, rpntr, (nbr + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(d_cp, cpntr, (nbc + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    HIPOK(hipEventCreate(&e0));
    HIPOK(hipEventCreate(&e1));
    HIPOK(hipEventRecord(e0, nullptr));
    auto g256 = [](int64_t n) { return dim3((unsigned)((n + 255) / 256)); };
    hipLaunchKernelGGL(sb_part_map, g256(nrows), dim3(256), 0, nullptr, d_rp, nbr, nrows, d_rmap);
    hipLaunchKernelGGL(sb_part_map, g256(ncols), dim3(256), 0, nullptr, d_cp, nbc, ncols, d_cmap);
    hipLaunchKernelGGL(sb_entry_cols, dim3((unsigned)((ncols + 3) / 4)), dim3(256), 0, nullptr, d_ip, ncols, d_col);
    HIPOK(hipMemsetAsync(d_bits, 0, (int64_t)nbr * rw * sizeof(uint32_t), nullptr));
    if (nnz > 0)
        hipLaunchKernelGGL(sb_mark_blocks, g256(nnz), dim3(256), 0, nullptr, d_ix, d_col, nnz, d_rmap, d_cmap, rw,
                           d_bits);
    hipLaunchKernelGGL(sb_row_prefix, g256(nbr), dim3(256), 0, nullptr, d_bits, nbr, rw, d_wpre, d_rowcnt);
    HIPOK(hipGetLastError());
    std::vector<int32_t> rowcnt(nbr);
    HIPOK(hipMemcpy(rowcnt.data(), d_rowcnt, nbr * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int64_t> rowbeg(nbr);
    h->bpntrb.resize(nbr);
    h->bpntre.resize(nbr);
    int64_t nb = 0;
    for (int32_t a = 0; a < nbr; ++a) {
        rowbeg[a] = nb;
        h->bpntrb[a] = rowcnt[a] ? nb : -1;
        nb += rowcnt[a];
        h->bpntre[a] = rowcnt[a] ? nb : -1;
    }
    h->nblocks = nb;
    HIPOK(hipMemcpy(d_rowbeg, rowbeg.data(), nbr * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPOK(hipMalloc((void **)&h->d_bindx, std::max<int64_t>(nb, 1) * sizeof(int64_t)));
    HIPOK(hipMalloc((void **)&h->d_indx, (nb + 1) * sizeof(int64_t)));
    HIPOK(hipMalloc((void **)&h->d_bnnz, std::max<int64_t>(nb, 1) * sizeof(int32_t)));
    CHECKRC(b.alloc((void **)&d_size, nb * sizeof(int64_t)));
    int64_t nv = 0;
    if (nb > 0) {
        hipLaunchKernelGGL(sb_kept_blocks, g256((int64_t)nbr * rw), dim3(256), 0, nullptr, d_bits, d_wpre, d_rowbeg,
                           nbr, rw, d_rp, d_cp, h->d_bindx, d_size);
        CHECKRC(scan_exclusive(d_size, nb, h->d_indx, b));
        int64_t last[2] = {0, 0};
        HIPOK(hipMemcpy(&last[0], h->d_indx + nb - 1, sizeof(int64_t), hipMemcpyDeviceToHost));
        HIPOK(hipMemcpy(&last[1], d_size + nb - 1, sizeof(int64_t), hipMemcpyDeviceToHost));
        nv = last[0] + last[1];
    }
    HIPOK(hipMemcpy(h->d_indx + nb, &nv, sizeof(int64_t), hipMemcpyHostToDevice));
    h->nval = nv;
    HIPOK(hipMalloc((void **)&h->d_val, std::max<int64_t>(nv, 1) * sizeof(double)));
    HIPOK(hipMemsetAsync(h->d_val, 0, std::max<int64_t>(nv, 1) * sizeof(double), nullptr));
    HIPOK(hipMemsetAsync(h->d_bnnz, 0, std::max<int64_t>(nb, 1) * sizeof(int32_t), nullptr));
    if (nnz > 0)
        hipLaunchKernelGGL(sb_scatter_values, g256(nnz), dim3(256), 0, nullptr, d_ix, d_col, d_v, nnz, d_rmap,
                           d_cmap, d_bits, d_wpre, d_rowbeg, rw, d_rp, d_cp, h->d_indx, h->d_val, h->d_bnnz);
    HIPOK(hipEventRecord(e1, nullptr));
    const hipError_t le = hipGetLastError();
    const hipError_t se = hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    HIPOK(le);
    HIPOK(se);
    if (device_ms) *device_ms = ms;
    if (nblocks) *nblocks = h->nblocks;
    if (nval) *nval = h->nval;
    *out = h.release();
    return SABLE_OK;
}

int sable_inspect_fetch(SableInspect *h, double *val, int64_t *indx, int64_t *bindx, int64_t *bpntrb,
                        int64_t *bpntre, int32_t *block_nnz)
{
    if (!h) return fail(SABLE_ERR_INVALID, "null inspect handle");
    if (val && h->nval) HIPOK(hipMemcpy(val, h->d_val, h->nval * sizeof(double), hipMemcpyDeviceToHost));
    if (indx) HIPOK(hipMemcpy(indx, h->d_indx, (h->nblocks + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (bindx && h->nblocks) HIPOK(hipMemcpy(bindx, h->d_bindx, h->nblocks * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (block_nnz && h->nblocks)
        HIPOK(hipMemcpy(block_nnz, h->d_bnnz, h->nblocks * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (bpntrb) std::copy(h->bpntrb.begin(), h->bpntrb.end(), bpntrb);
    if (bpntre) std::copy(h->bpntre.begin(), h->bpntre.end(), bpntre);
    return SABLE_OK;
}

int sable_inspect_destroy(SableInspect *h)
{
    delete h;
    return SABLE_OK;
}

// ---- compressed (.vbrc) arrays for given routes

struct SableCompress {
    int64_t nval2 = 0, nub = 0, nnz = 0;   // nnz: CSR entries including the padding entry
    bool has_csr = false;
    std::vector<int64_t> indx2, ublocks;
    double *d_val2 = nullptr, *d_vals = nullptr;
    int64_t *d_rowptr = nullptr, *d_cols = nullptr;
    int64_t nrows = 0;
    ~SableCompress()
    {
        hipFree(d_val2);
        hipFree(d_vals);
        hipFree(d_rowptr);
        hipFree(d_cols);
    }
};

}  // extern "C"

namespace {

// Compression of VBR arrays already on the device (val, indx, bindx) for
// the given routes; indx, the partitions and the block-row pointers on the
// host.  Shared by sable_compress_create (after uploading) and
// sable_inspect_compress (the inspector's own arrays).
int compress_on_device(const double *d_val, const int64_t *d_indx, const int64_t *d_bindx, const int64_t *indx,
                       int64_t nblocks, const int64_t *rpntr, int32_t nbr, const int64_t *cpntr, int32_t nbc,
                       const int64_t *bpntrb, const int64_t *bpntre, const uint8_t *routes, SableCompress **out,
                       int64_t *nval2, int64_t *nublocks, int64_t *csr_nnz, float *device_ms)
{
    const int64_t nrows = rpntr[nbr], ncols = cpntr[nbc];
    std::unique_ptr<SableCompress> h(new SableCompress());
    h->nrows = nrows;
    std::vector<int64_t> dblk;
    std::vector<uint8_t> sp(std::max<int64_t>(nblocks, 1), 0);
    h->indx2.push_back(0);
    for (int64_t k = 0; k < nblocks; ++k) {
        if (routes[k] == SABLE_ROUTE_SPARSE) {
            sp[k] = 1;
            h->ublocks.push_back(k);
        } else {
            dblk.push_back(k);
            h->indx2.push_back(h->indx2.back() + indx[k + 1] - indx[k]);
        }
    }
    h->nval2 = h->indx2.back();
    h->nub = (int64_t)h->ublocks.size();
    DevBufs b;
    int64_t *d_rp = nullptr, *d_cp = nullptr, *d_bb = nullptr, *d_be = nullptr, *d_dblk = nullptr,
            *d_indx2 = nullptr, *d_cnt = nullptr;
    uint8_t *d_sp = nullptr;
    int32_t *d_rmap = nullptr, *d_last = nullptr;
    CHECKRC(b.alloc((void **)&d_rp, (nbr + 1) * sizeof(int64_t)));
    CHECKRC(b.alloc((void **)&d_cp, (nbc + 1) * sizeof(int64_t)));
    CHECKRC(b.alloc((void **)&d_bb, nbr * sizeof(int64_t)));
    CHECKRC(b.alloc((void **)&d_be, nbr * sizeof(int64_t)));
    CHECKRC(b.alloc((void **)&d_dblk, dblk.size() * sizeof(int64_t)));
    CHECKRC(b.alloc((void **)&d_indx2, h->indx2.size() * sizeof(int64_t)));
    CHECKRC(b.alloc((void **)&d_cnt, nrows * sizeof(int64_t)));
    CHECKRC(b.alloc((void **)&d_sp, sp.size()));
    CHECKRC(b.alloc((void **)&d_rmap, nrows * sizeof(int32_t)));
    CHECKRC(b.alloc((void **)&d_last, sizeof(int32_t)));
    HIPOK(hipMemcpy(d_rp, rpntr, (nbr + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(d_cp, cpntr, (nbc + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(d_bb, bpntrb, nbr * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(d_be, bpntre, nbr * sizeof(int64_t), hipMemcpyHostToDevice));
    if (!dblk.empty()) HIPOK(hipMemcpy(d_dblk, dblk.data(), dblk.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(d_indx2, h->indx2.data(), h->indx2.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(d_sp, sp.data(), sp.size(), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    HIPOK(hipEventCreate(&e0));
    HIPOK(hipEventCreate(&e1));
    HIPOK(hipEventRecord(e0, nullptr));
    auto g256 = [](int64_t n) { return dim3((unsigned)((n + 255) / 256)); };
    HIPOK(hipMalloc((void **)&h->d_val2, std::max<int64_t>(h->nval2, 1) * sizeof(double)));
    if (!dblk.empty())
        hipLaunchKernelGGL(sb_gather_dense, dim3((unsigned)((dblk.size() + 3) / 4)), dim3(256), 0, nullptr, d_val,
                           d_indx, d_dblk, d_indx2, (int64_t)dblk.size(), h->d_val2);
    int64_t nnz = 0;
    if (h->nub > 0) {
        hipLaunchKernelGGL(sb_part_map, g256(nrows), dim3(256), 0, nullptr, d_rp, nbr, nrows, d_rmap);
        HIPOK(hipMemsetAsync(d_last, 0, sizeof(int32_t), nullptr));
        hipLaunchKernelGGL(sb_csr_rows, g256(nrows), dim3(256), 0, nullptr, d_val, d_indx, d_bindx, d_rp, d_cp,
                           d_bb, d_be, d_sp, d_rmap, nrows, ncols, d_cnt, nullptr, nullptr, nullptr, d_last);
        HIPOK(hipMalloc((void **)&h->d_rowptr, (nrows + 1) * sizeof(int64_t)));
        CHECKRC(scan_exclusive(d_cnt, nrows, h->d_rowptr, b));
        int64_t lastptr = 0, lastcnt = 0;
        int32_t lastcol = 0;
        HIPOK(hipMemcpy(&lastptr, h->d_rowptr + nrows - 1, sizeof(int64_t), hipMemcpyDeviceToHost));
        HIPOK(hipMemcpy(&lastcnt, d_cnt + nrows - 1, sizeof(int64_t), hipMemcpyDeviceToHost));
        HIPOK(hipMemcpy(&lastcol, d_last, sizeof(int32_t), hipMemcpyDeviceToHost));
        nnz = lastptr + lastcnt;
        if (nnz > 0) {
            // the reference's padding entry (utils/convert_real_to_vbr.py:75-80)
            const bool pad = !(lastcnt > 0 && lastcol);
            h->has_csr = true;
            h->nnz = nnz + (pad ? 1 : 0);
            HIPOK(hipMalloc((void **)&h->d_cols, h->nnz * sizeof(int64_t)));
            HIPOK(hipMalloc((void **)&h->d_vals, h->nnz * sizeof(double)));
            hipLaunchKernelGGL(sb_csr_rows, g256(nrows), dim3(256), 0, nullptr, d_val, d_indx, d_bindx, d_rp, d_cp,
                               d_bb, d_be, d_sp, d_rmap, nrows, ncols, d_cnt, h->d_rowptr, h->d_cols, h->d_vals,
                               d_last);
            if (pad) {
                const int64_t c = ncols - 1;
                const double z = 0.0;
                HIPOK(hipMemcpy(h->d_cols + nnz, &c, sizeof(int64_t), hipMemcpyHostToDevice));
                HIPOK(hipMemcpy(h->d_vals + nnz, &z, sizeof(double), hipMemcpyHostToDevice));
            }
            HIPOK(hipMemcpy(h->d_rowptr + nrows, &h->nnz, sizeof(int64_t), hipMemcpyHostToDevice));
        }
    }
    HIPOK(hipEventRecord(e1, nullptr));
    const hipError_t le = hipGetLastError();
    const hipError_t se = hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    HIPOK(le);
    HIPOK(se);
    if (device_ms) *device_ms = ms;
    if (nval2) *nval2 = h->nval2;
    if (nublocks) *nublocks = h->nub;
    if (csr_nnz) *csr_nnz = h->has_csr ? h->nnz : -1;
    *out = h.release();
    return SABLE_OK;
}

}  // namespace

extern "C" {

int sable_compress_create(const double *val, int64_t nval, const int64_t *indx, const int64_t *bindx,
                          int64_t nblocks, const int64_t *rpntr, int32_t nbr, const int64_t *cpntr, int32_t nbc,
                          const int64_t *bpntrb, const int64_t *bpntre, const uint8_t *routes,
                          SableCompress **out, int64_t *nval2, int64_t *nublocks, int64_t *csr_nnz,
                          float *device_ms)
{
    if (g_device < 0) return fail(SABLE_ERR_NO_DEVICE, "sable_init not called");
    if (!out || !indx || !bpntrb || !bpntre || nblocks < 0 || nval < 0) return fail(SABLE_ERR_INVALID, "bad compress arguments");
    *out = nullptr;
    if (nblocks > 0 && (!val || !bindx || !routes)) return fail(SABLE_ERR_INVALID, "null val/bindx/routes");
    const int64_t nrows = rpntr ? rpntr[nbr] : 0, ncols = cpntr ? cpntr[nbc] : 0;
    CHECKRC(check_partition(rpntr, nbr, nrows, "rpntr"));
    CHECKRC(check_partition(cpntr, nbc, ncols, "cpntr"));
    // the kernels read block k only inside [indx[k], indx[k] + h w): check it
    if (indx[0] != 0 || indx[nblocks] != nval) return fail(SABLE_ERR_INVALID, "indx must run from 0 to len(val)");
    int64_t next = 0;
    for (int32_t a = 0; a < nbr; ++a) {
        if (bpntrb[a] == -1 && bpntre[a] == -1) continue;
        if (bpntrb[a] != next || bpntre[a] < bpntrb[a] || bpntre[a] > nblocks)
            return fail(SABLE_ERR_INVALID, "bpntrb/bpntre must lay the block rows out in order");
        const int64_t h = rpntr[a + 1] - rpntr[a];
        for (int64_t k = bpntrb[a]; k < bpntre[a]; ++k) {
            if (bindx[k] < 0 || bindx[k] >= nbc) return fail(SABLE_ERR_INVALID, "bindx out of range");
            if (indx[k + 1] - indx[k] != h * (cpntr[bindx[k] + 1] - cpntr[bindx[k]]))
                return fail(SABLE_ERR_INVALID, "indx step must equal the block size");
        }
        next = bpntre[a];
    }
    if (next != nblocks) return fail(SABLE_ERR_INVALID, "bpntre must end at the block count");
    DevBufs b;
    double *d_val = nullptr;
    int64_t *d_indx = nullptr, *d_bindx = nullptr;
    CHECKRC(b.alloc((void **)&d_val, nval * sizeof(double)));
    CHECKRC(b.alloc((void **)&d_indx, (nblocks + 1) * sizeof(int64_t)));
    CHECKRC(b.alloc((void **)&d_bindx, nblocks * sizeof(int64_t)));
    if (nval) HIPOK(hipMemcpy(d_val, val, nval * sizeof(double), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(d_indx, indx, (nblocks + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    if (nblocks) HIPOK(hipMemcpy(d_bindx, bindx, nblocks * sizeof(int64_t), hipMemcpyHostToDevice));
    return compress_on_device(d_val, d_indx, d_bindx, indx, nblocks, rpntr, nbr, cpntr, nbc, bpntrb, bpntre, routes,
                              out, nval2, nublocks, csr_nnz, device_ms);
}

int sable_compress_fetch(SableCompress *h, double *val2, int64_t *indx2, int64_t *ublocks, int64_t *indptr,
                         int64_t *indices, double *csr_val)
{
    if (!h) return fail(SABLE_ERR_INVALID, "null compress handle");
    if (val2 && h->nval2) HIPOK(hipMemcpy(val2, h->d_val2, h->nval2 * sizeof(double), hipMemcpyDeviceToHost));
    if (indx2) std::copy(h->indx2.begin(), h->indx2.end(), indx2);
    if (ublocks) std::copy(h->ublocks.begin(), h->ublocks.end(), ublocks);
    if (h->has_csr) {
        if (indptr) HIPOK(hipMemcpy(indptr, h->d_rowptr, (h->nrows + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
        if (indices) HIPOK(hipMemcpy(indices, h->d_cols, h->nnz * sizeof(int64_t), hipMemcpyDeviceToHost));
        if (csr_val) HIPOK(hipMemcpy(csr_val, h->d_vals, h->nnz * sizeof(double), hipMemcpyDeviceToHost));
    }
    return SABLE_OK;
}

int sable_compress_destroy(SableCompress *h)
{
    delete h;
    return SABLE_OK;
}

int sable_inspect_compress(SableInspect *ih, const uint8_t *routes, SableCompress **out, int64_t *nval2,
                           int64_t *nublocks, int64_t *csr_nnz, float *device_ms)
{
    if (g_device < 0) return fail(SABLE_ERR_NO_DEVICE, "sable_init not called");
    if (!ih || !out || (ih->nblocks > 0 && !routes)) return fail(SABLE_ERR_INVALID, "bad inspect_compress arguments");
    *out = nullptr;
    std::vector<int64_t> indx(ih->nblocks + 1);
    HIPOK(hipMemcpy(indx.data(), ih->d_indx, indx.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    return compress_on_device(ih->d_val, ih->d_indx, ih->d_bindx, indx.data(), ih->nblocks, ih->rpntr.data(),
                              (int32_t)ih->rpntr.size() - 1, ih->cpntr.data(), (int32_t)ih->cpntr.size() - 1,
                              ih->bpntrb.data(), ih->bpntre.data(), routes, out, nval2, nublocks, csr_nnz,
                              device_ms);
}

// ---- device memory helpers

int sable_malloc(void **ptr, int64_t bytes)
{
    if (!ptr || bytes < 0) return fail(SABLE_ERR_INVALID, "bad malloc arguments");
    HIPOK(hipMalloc(ptr, std::max<int64_t>(bytes, 1)));
    return SABLE_OK;
}

int sable_free(void *ptr)
{
    if (ptr) HIPOK(hipFree(ptr));
    return SABLE_OK;
}

int sable_memcpy_h2d(void *dst, const void *src, int64_t bytes)
{
    HIPOK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SABLE_OK;
}

int sable_memcpy_d2h(void *dst, const void *src, int64_t bytes)
{
    HIPOK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return SABLE_OK;
}

int sable_synchronize(void)
{
    HIPOK(hipDeviceSynchronize());
    return SABLE_OK;
}

}  // extern "C"
