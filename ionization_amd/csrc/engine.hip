// Native runtime of the mesh-TDSE library: context, device memory pool,
// ragged work tables, hipGraph-captured steps and the C ABI of tdse_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tdse_hip.h"
#include "common.h"
#include "cyl_kernels.h"
#include "line_kernels.h"
#include "sh_kernels.h"
#include "tdma_kernels.h"

namespace {

constexpr int kAbiVersion = 1;
constexpr size_t kAlign = 256;
// Steps captured in the longer of the two step graphs (runs of at least this
// many steps replay it; the rest replay the one-step graph): one launch per 8
// steps took small batches 6-9 % faster (one 200 x 1000 SH simulation: 0.047 ->
// 0.043 ms/step) and changed nothing for the 256-simulation scan.
constexpr int kGraphSteps = 8;
constexpr int kMaxL = 4096;          // largest l_bound accepted (c_l table size)
constexpr int kSlotCapacity = 8192;  // simulations of each kind per context
constexpr int kSideStreams = 3;      // extra streams for concurrent R-phase buckets (default priority)
constexpr int kMaxPairThreads = 512; // largest workgroup of the two-line (velocity odd l-pair) R-phase
constexpr int kNoFlags = 0;          // health flags of a freshly uploaded state
static_assert(TDSE_FLAG_NONFINITE == TDSE_HEALTH_NONFINITE && TDSE_FLAG_NORM_GROWTH == TDSE_HEALTH_NORM_GROWTH,
              "health flag values of the ABI and the kernels differ");

thread_local std::string g_lib_error;

// First-fit sub-allocator over one device buffer reserved at context creation.
class Arena {
   public:
    char* base = nullptr;
    size_t size = 0;
    size_t used = 0;

    bool init(size_t bytes) {
        size = (bytes + kAlign - 1) / kAlign * kAlign;
        if (hipMalloc(&base, size) != hipSuccess) {
            base = nullptr;
            return false;
        }
        free_[0] = size;
        return true;
    }
    void release() {
        if (base) (void)hipFree(base);
        base = nullptr;
    }
    // Returns nullptr when no free block is large enough.
    void* alloc(size_t bytes) {
        bytes = std::max<size_t>(kAlign, (bytes + kAlign - 1) / kAlign * kAlign);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first;
                const size_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used_[off] = bytes;
                used += bytes;
                return base + off;
            }
        }
        return nullptr;
    }
    void dealloc(void* p) {
        if (!p) return;
        const size_t off = static_cast<char*>(p) - base;
        auto it = used_.find(off);
        if (it == used_.end()) return;
        size_t sz = it->second;
        used -= sz;
        used_.erase(it);
        size_t start = off;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && next->first == off + sz) {
            sz += next->second;
            free_.erase(next);
        }
        auto prev = free_.lower_bound(off);
        if (prev != free_.begin()) {
            --prev;
            if (prev->first + prev->second == off) {
                start = prev->first;
                sz += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = sz;
    }

   private:
    std::map<size_t, size_t> free_;
    std::map<size_t, size_t> used_;
};

struct ShHost {
    bool live = false;
    bool vel = false;    // velocity gauge (one more table row, h2 factors around the R-phase)
    bool adi = false;    // AlternatingDirectionImplicit (its own r and l sweeps instead of the R-phase)
    ShSim dev{};
    int n_data = 0;
    int wave_first = 0;  // this simulation's range in the wave table
    int wave_count = 0;
    std::vector<void*> blocks;
};

struct LineHost {
    bool live = false;
    bool complex_steps = false;  // some w_n has a nonzero real part
    LineSim dev{};
    int n_data = 0;
    std::vector<void*> blocks;
};

struct CylHost {
    bool live = false;
    CylSim dev{};
    int n_data = 0;
    std::vector<void*> blocks;
};

// Device table with host-side growth.
template <class T>
struct DevVec {
    T* ptr = nullptr;
    size_t cap = 0;
    size_t count = 0;
};

struct Bucket {
    int M = 0;  // exact rows per chunk-thread
    int P = 0;
    bool imag = false;  // Line: complex time steps (imaginary-time relaxation)
    int method = 0;     // Line: LINE_ADI, LINE_SIO_LEN or LINE_SIO_VEL
    int lanes = 1;      // cylindrical rho-sweep: lines per workgroup
    DevVec<int2> lines;  // SH: (slot, l); Line: (slot, 0); Cyl z: (slot, rho); Cyl rho: (slot, position group)
    DevVec<int2> vpairs[2];  // SH: the velocity-gauge l-pairs (slot, l) of even / odd l among these simulations
    DevVec<int2> fpairs;     // SH: velocity-gauge odd l-pairs solved with their h2 factors (k_sh_rphase_vpair)
};

void free_bucket(Bucket& b) {
    if (b.lines.ptr) (void)hipFree(b.lines.ptr);
    for (auto& v : b.vpairs)
        if (v.ptr) (void)hipFree(v.ptr);
    if (b.fpairs.ptr) (void)hipFree(b.fpairs.ptr);
    b = Bucket{};
}

// l segments of at most kSegRows rows per L-phase wave, balanced in size.
constexpr int kSegRows = kShSegRows;
int sh_segments(int L) { return (L + kSegRows - 1) / kSegRows; }

int n_data_points(int N, int every) {
    if (every >= 1) return N / every + 1 + ((N % every) != 0 ? 1 : 0);
    return 2;
}

}  // namespace

struct tdse_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Arena arena;
    std::string err;
    std::vector<ShHost> sh;
    std::vector<LineHost> line;
    std::vector<CylHost> cyl;
    ShSim* d_sh = nullptr;
    LineSim* d_line = nullptr;
    CylSim* d_cyl = nullptr;
    DevVec<int2> cyl_jobs;
    std::vector<Bucket> cyl_z_buckets;
    std::vector<Bucket> cyl_r_buckets;
    double* d_cl = nullptr;
    double* d_clv = nullptr;  // c_l (l + 1): the velocity gauge's h1 coupling per l
    int* d_ctl = nullptr;     // [1] nonzero while the next replay of the current run follows (ADI run-ahead)
    DevVec<ShWave> waves;     // length-gauge simulations' waves first, then velocity-gauge ones
    int n_len_waves = 0;
    DevVec<int2> jobs;
    std::vector<Bucket> sh_buckets;
    std::vector<int> sh_order;  // sh_buckets by decreasing work
    std::vector<Bucket> sh_adi_buckets;  // AlternatingDirectionImplicit l-lines (slot, l) by (rows, chunks)
    DevVec<int2> adi_waves;              // AlternatingDirectionImplicit positions: (slot, first position of a wave)
    std::vector<Bucket> sh_adi_lbuckets;  // ADI position groups (slot, group) of the partition l-sweep, by rows
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_join[3] = {nullptr, nullptr, nullptr};
    std::vector<Bucket> line_buckets;
    bool dirty = true;
    bool sh_norm_by_l = false;   // some live SH simulation reduces the norm in every l (L-phase instance)
    bool cyl_pending = false;    // a cylindrical simulation was added or uploaded: it starts from an exact state
    bool graph_cyl_pre = false;  // the captured step includes the cylindrical pre-sweep
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph_k = nullptr;  // kGraphSteps steps in one graph (fewer launches for small batches)
    hipGraphExec_t exec_k = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    double2* scratch = nullptr;  // this device's copy of `scratch` (tdse_tdma_batched_noalloc)
    double last_run_ms = 0.0;
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    // (kind, start event index); kinds: SH L-phase, SH R-phase, Cyl z, Cyl rho, SH velocity-gauge h2
    // factors, SH velocity-gauge odd l-pair R-phase
    std::vector<std::pair<int, int>> ev_pairs;
    double prof_ms[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int64_t prof_n[6] = {0, 0, 0, 0, 0, 0};
};

namespace {

int fail(tdse_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    else g_lib_error = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                          \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail((ctx), TDSE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

template <class T>
int devvec_upload(tdse_ctx* ctx, DevVec<T>& v, const std::vector<T>& host) {
    if (host.size() > v.cap) {
        if (v.ptr) HIP_TRY(ctx, hipFree(v.ptr));
        v.cap = std::max<size_t>(host.size(), 2 * v.cap);
        HIP_TRY(ctx, hipMalloc(&v.ptr, v.cap * sizeof(T)));
    }
    v.count = host.size();
    if (!host.empty())
        HIP_TRY(ctx, hipMemcpyAsync(v.ptr, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return TDSE_OK;
}

void drop_graph(tdse_ctx* ctx) {
    if (ctx->exec) (void)hipGraphExecDestroy(ctx->exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    if (ctx->exec_k) (void)hipGraphExecDestroy(ctx->exec_k);
    if (ctx->graph_k) (void)hipGraphDestroy(ctx->graph_k);
    ctx->exec = nullptr;
    ctx->graph = nullptr;
    ctx->exec_k = nullptr;
    ctx->graph_k = nullptr;
}

// Rebuild the ragged work tables from the live simulations.
int rebuild_tables(tdse_ctx* ctx) {
    std::vector<ShWave> waves;
    std::vector<int2> jobs;
    ctx->n_len_waves = -1;
    std::map<std::pair<int, int>, std::vector<int2>> shb, lnb;  // (rows, chunks [+ 4096 if complex steps])
    std::map<std::pair<int, int>, std::vector<int2>> vpb[2];    // velocity l-pairs by parity, (rows, chunks)
    std::map<std::pair<int, int>, std::vector<int2>> fpb;       // velocity odd l-pairs solved in one workgroup
    std::map<std::pair<int, int>, std::vector<int2>> adb;       // ADI l-lines, (rows, chunks)
    std::vector<int2> adw;                                      // ADI position waves (Thomas l-sweep)
    std::map<std::pair<int, int>, std::vector<int2>> alb;       // ADI position groups (partition l-sweep), (rows, 0)
    ctx->sh_norm_by_l = false;
    for (int pass = 0; pass < 2; ++pass)
    for (int s = 0; s < (int)ctx->sh.size(); ++s) {
        ShHost& h = ctx->sh[s];
        if (!h.live || (int)h.vel != pass) continue;
        if (pass == 1 && ctx->n_len_waves < 0) ctx->n_len_waves = (int)waves.size();
        h.wave_first = (int)waves.size();
        const int nseg = sh_segments(h.dev.L);
        const int seg = (h.dev.L + nseg - 1) / nseg;
        const int wpl = h.dev.npad / TDSE_WAVE;
        for (int sg = 0; sg < nseg; ++sg) {
            const int lo = sg * seg, hi = std::min(h.dev.L, lo + seg);
            for (int w = 0; w < wpl; ++w) waves.push_back(ShWave{s, w * TDSE_WAVE, lo, hi, sg * wpl + w});
        }
        h.wave_count = (int)waves.size() - h.wave_first;
        jobs.push_back(make_int2(s, -1));
        if (h.dev.norm_by_l) {
            jobs.push_back(make_int2(s, -2));
            ctx->sh_norm_by_l = true;
        }
        for (int b = 0; b < h.dev.n_basis; ++b) jobs.push_back(make_int2(s, b));
        if (h.adi) {
            for (int l = 0; l < h.dev.L; ++l) adb[{h.dev.m, h.dev.P}].push_back(make_int2(s, l));
            // l-direction: the partition solver where kAdiChunks chunks of at most 16 rows cover l_bound
            const int ml = 2 * ((h.dev.L + 2 * kAdiChunks - 1) / (2 * kAdiChunks));
            if (ml <= 16) {
                for (int gq = 0; gq < h.dev.npad / kAdiLanes; ++gq) alb[{std::max(ml, 2), 0}].push_back(make_int2(s, gq));
            } else {
                for (int w = 0; w < wpl; ++w) adw.push_back(make_int2(s, w * TDSE_WAVE));
            }
            continue;
        }
        auto& v = shb[{h.dev.m, h.dev.P}];
        // velocity gauge: where two lines fit one workgroup (2 P <= 512) the odd l-pairs are
        // solved together with their h2 factors; the rows outside them stay single lines
        const bool fuse = h.vel && 2 * h.dev.P <= kMaxPairThreads;
        for (int l = 0; l < h.dev.L; ++l) {
            const bool in_odd_pair = (l & 1) ? (l + 1 < h.dev.L) : (l >= 1);
            if (!(fuse && in_odd_pair)) v.push_back(make_int2(s, l));
        }
        if (h.vel)
            for (int l = 0; l + 1 < h.dev.L; ++l) {
                if (fuse && (l & 1)) fpb[{h.dev.m, h.dev.P}].push_back(make_int2(s, l));
                else vpb[l & 1][{h.dev.m, h.dev.P}].push_back(make_int2(s, l));
            }
    }
    if (ctx->n_len_waves < 0) ctx->n_len_waves = (int)waves.size();
    for (int s = 0; s < (int)ctx->line.size(); ++s) {
        const LineHost& h = ctx->line[s];
        if (!h.live) continue;
        lnb[{h.dev.m, h.dev.P + (h.complex_steps ? 4096 : 0) + 8192 * h.dev.method}].push_back(make_int2(s, 0));
    }
    std::vector<int2> cjobs;
    std::map<std::pair<int, int>, std::vector<int2>> czb;             // (rows, chunks)
    std::map<std::pair<int, int>, std::vector<int2>> crb;             // (rows + 64 * lanes, chunks)
    for (int s = 0; s < (int)ctx->cyl.size(); ++s) {
        const CylHost& h = ctx->cyl[s];
        if (!h.live) continue;
        const CylSim& C = h.dev;
        cjobs.push_back(make_int2(s, -1));
        for (int b = 0; b < C.n_basis; ++b) cjobs.push_back(make_int2(s, b));
        auto& vz = czb[{C.Mz, C.Pz}];
        for (int q = 0; 2 * q < C.nr; ++q) vz.push_back(make_int2(s, q));  // pairs of rho-lines
        auto& vr = crb[{C.Mr + 64 * C.lanes, C.Cr}];
        for (int g = 0; g < C.n_groups; ++g) vr.push_back(make_int2(s, g));
    }
    int rc;
    if ((rc = devvec_upload(ctx, ctx->waves, waves))) return rc;
    if ((rc = devvec_upload(ctx, ctx->jobs, jobs))) return rc;
    if ((rc = devvec_upload(ctx, ctx->cyl_jobs, cjobs))) return rc;
    auto fill = [&](std::vector<Bucket>& dst, std::map<std::pair<int, int>, std::vector<int2>>& src) -> int {
        for (auto& b : dst) free_bucket(b);
        dst.clear();
        for (auto& kv : src) {
            Bucket b;
            b.M = kv.first.first;
            b.P = kv.first.second % 4096;
            b.imag = (kv.first.second % 8192) >= 4096;
            b.method = kv.first.second / 8192;
            int r = devvec_upload(ctx, b.lines, kv.second);
            if (r) return r;
            dst.push_back(b);
        }
        return TDSE_OK;
    };
    for (auto& kv : fpb) shb[kv.first];  // a bucket for every layout with fused pairs, even without single lines
    if ((rc = fill(ctx->sh_buckets, shb))) return rc;
    if ((rc = fill(ctx->sh_adi_buckets, adb))) return rc;
    if ((rc = fill(ctx->sh_adi_lbuckets, alb))) return rc;
    if ((rc = devvec_upload(ctx, ctx->adi_waves, adw))) return rc;
    for (auto& b : ctx->sh_buckets) {
        for (int par = 0; par < 2; ++par) {
            auto it = vpb[par].find({b.M, b.P});
            if (it != vpb[par].end() && (rc = devvec_upload(ctx, b.vpairs[par], it->second))) return rc;
        }
        auto it = fpb.find({b.M, b.P});
        if (it != fpb.end() && (rc = devvec_upload(ctx, b.fpairs, it->second))) return rc;
    }
    if ((rc = fill(ctx->line_buckets, lnb))) return rc;
    if ((rc = fill(ctx->cyl_z_buckets, czb))) return rc;
    if ((rc = fill(ctx->cyl_r_buckets, crb))) return rc;
    {
        std::vector<std::pair<double, int>> w;
        for (size_t i = 0; i < ctx->sh_buckets.size(); ++i) {
            // rows swept per step: the R-phase's lines, and the velocity-gauge h2 factors' rows
            // (two rows per pair, before and after the solve)
            const Bucket& b = ctx->sh_buckets[i];
            const double rows = (double)b.lines.count + 2.0 * (double)b.fpairs.count +
                                4.0 * (double)(b.vpairs[0].count + b.vpairs[1].count);
            w.push_back({-rows * b.P * b.M, (int)i});
        }
        std::sort(w.begin(), w.end());
        ctx->sh_order.clear();
        for (auto& pr : w) ctx->sh_order.push_back(pr.second);
    }
    for (auto& b : ctx->cyl_r_buckets) {  // key first = rows + 64 * lanes
        b.lanes = b.M / 64;
        b.M %= 64;
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    ctx->dirty = false;
    return TDSE_OK;
}

hipEvent_t next_event(tdse_ctx* ctx) {
    if (ctx->ev_used == ctx->ev_pool.size()) {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        ctx->ev_pool.push_back(e);
    }
    return ctx->ev_pool[ctx->ev_used++];
}

void prof_begin(tdse_ctx* ctx, int kind) {
    if (!ctx->profiling) return;
    ctx->ev_pairs.push_back({kind, (int)ctx->ev_used});
    (void)hipEventRecord(next_event(ctx), ctx->stream);
}
void prof_end(tdse_ctx* ctx) {
    if (!ctx->profiling) return;
    (void)hipEventRecord(next_event(ctx), ctx->stream);
}

// Form of a CN line kernel with M rows per thread and P threads: the
// explicit right-hand side for M >= 10; for M <= 8 the Cayley form, keeping
// the original rows in LDS when that fits the 64 KiB a workgroup gets by
// default, else re-reading them. Chosen from kernel-trace timings.
int cn_form(int M, int P) {
    if (M >= 10) return CN_EXPLICIT;
    return (3 + M) * (size_t)P * sizeof(double2) <= 65536 ? CN_CAYLEY_LDS : CN_CAYLEY_MEM;
}
size_t cn_lds_bytes(int M, int P, int form) {
    return (size_t)(form == CN_CAYLEY_LDS ? 3 + M : 3) * P * sizeof(double2);
}

// Launch shapes per rows-per-thread M (block size, minimum waves per EU).
template <int M, int MAXT, int WPE>
void launch_rphase(tdse_ctx* ctx, const Bucket& b, hipStream_t st) {
    const int form = cn_form(M, b.P);
    const size_t lds = cn_lds_bytes(M, b.P, form);
    const dim3 grid((unsigned)b.lines.count), block(b.P);
    if (form == CN_EXPLICIT)
        hipLaunchKernelGGL((k_sh_rphase<M, MAXT, WPE, (M >= 10 ? CN_EXPLICIT : CN_CAYLEY_MEM)>), grid, block, lds,
                           st, ctx->d_sh, b.lines.ptr, (int)b.lines.count);
    else if (form == CN_CAYLEY_LDS)
        hipLaunchKernelGGL((k_sh_rphase<M, MAXT, WPE, (M >= 10 ? CN_EXPLICIT : CN_CAYLEY_LDS)>), grid, block, lds,
                           st, ctx->d_sh, b.lines.ptr, (int)b.lines.count);
    else
        hipLaunchKernelGGL((k_sh_rphase<M, MAXT, WPE, (M >= 10 ? CN_EXPLICIT : CN_CAYLEY_MEM)>), grid, block, lds,
                           st, ctx->d_sh, b.lines.ptr, (int)b.lines.count);
}

void dispatch_rphase(tdse_ctx* ctx, const Bucket& b, hipStream_t st) {
    switch (b.M) {
        case 2: launch_rphase<2, 512, 4>(ctx, b, st); break;
        case 4: launch_rphase<4, 512, 4>(ctx, b, st); break;
        case 6: launch_rphase<6, 512, 4>(ctx, b, st); break;
        case 8: launch_rphase<8, 512, 4>(ctx, b, st); break;
        // radial lines past 4096 points: 512-thread workgroups of deep tiles
        case 10: b.P > 256 ? launch_rphase<10, 512, 1>(ctx, b, st) : launch_rphase<10, 256, 3>(ctx, b, st); break;
        case 12: b.P > 256 ? launch_rphase<12, 512, 1>(ctx, b, st) : launch_rphase<12, 256, 2>(ctx, b, st); break;
        case 14: b.P > 256 ? launch_rphase<14, 512, 1>(ctx, b, st) : launch_rphase<14, 256, 2>(ctx, b, st); break;
        default: b.P > 256 ? launch_rphase<16, 512, 1>(ctx, b, st) : launch_rphase<16, 256, 1>(ctx, b, st); break;
    }
}

template <int M, int MAXT, int WPE, bool EXPL>
void launch_adi_r(tdse_ctx* ctx, const Bucket& b, hipStream_t st) {
    const size_t lds = 3 * (size_t)b.P * sizeof(double2);
    hipLaunchKernelGGL((k_sh_adi_r<M, MAXT, WPE, EXPL>), dim3((unsigned)b.lines.count), dim3(b.P), lds, st,
                       ctx->d_sh, b.lines.ptr, (int)b.lines.count, ctx->d_ctl);
}

// the R-phase's launch shapes
template <bool EXPL>
void dispatch_adi_r(tdse_ctx* ctx, const Bucket& b, hipStream_t st) {
    switch (b.M) {
        case 2: launch_adi_r<2, 512, 4, EXPL>(ctx, b, st); break;
        case 4: launch_adi_r<4, 512, 4, EXPL>(ctx, b, st); break;
        case 6: launch_adi_r<6, 512, 4, EXPL>(ctx, b, st); break;
        case 8: launch_adi_r<8, 512, EXPL ? 4 : 3, EXPL>(ctx, b, st); break;  // the run-ahead's registers
        case 10: b.P > 256 ? launch_adi_r<10, 512, 1, EXPL>(ctx, b, st) : launch_adi_r<10, 256, 3, EXPL>(ctx, b, st); break;
        case 12: b.P > 256 ? launch_adi_r<12, 512, 1, EXPL>(ctx, b, st) : launch_adi_r<12, 256, 2, EXPL>(ctx, b, st); break;
        case 14: b.P > 256 ? launch_adi_r<14, 512, 1, EXPL>(ctx, b, st) : launch_adi_r<14, 256, 2, EXPL>(ctx, b, st); break;
        default: b.P > 256 ? launch_adi_r<16, 512, 1, EXPL>(ctx, b, st) : launch_adi_r<16, 256, 2, EXPL>(ctx, b, st); break;
    }
}

template <int M>
void launch_adi_lp(tdse_ctx* ctx, const Bucket& b, hipStream_t st) {
    const size_t lds = 3 * (size_t)kAdiChunks * kAdiLanes * sizeof(double2);
    hipLaunchKernelGGL(k_sh_adi_lp<M>, dim3((unsigned)b.lines.count), dim3(kAdiChunks * kAdiLanes), lds, st,
                       ctx->d_sh, b.lines.ptr, (int)b.lines.count, ctx->d_cl);
}

void dispatch_adi_lp(tdse_ctx* ctx, const Bucket& b, hipStream_t st) {
    switch (b.M) {
        case 2: launch_adi_lp<2>(ctx, b, st); break;
        case 4: launch_adi_lp<4>(ctx, b, st); break;
        case 6: launch_adi_lp<6>(ctx, b, st); break;
        case 8: launch_adi_lp<8>(ctx, b, st); break;
        case 10: launch_adi_lp<10>(ctx, b, st); break;
        case 12: launch_adi_lp<12>(ctx, b, st); break;
        case 14: launch_adi_lp<14>(ctx, b, st); break;
        default: launch_adi_lp<16>(ctx, b, st); break;
    }
}

template <int M, int WPE>
void launch_rphase_vpair(tdse_ctx* ctx, const Bucket& b, hipStream_t st) {
    // the reduced system (3 P) and alpha_i of rows 0 .. M - 3 (M - 2 P) per line, two lines
    const size_t lds = (size_t)(3 + M - 2) * b.P * 2 * sizeof(double2);
    hipLaunchKernelGGL((k_sh_rphase_vpair<M, kMaxPairThreads, WPE>), dim3((unsigned)b.fpairs.count), dim3(2 * b.P),
                       lds, st, ctx->d_sh, b.fpairs.ptr, (int)b.fpairs.count, ctx->d_cl);
}

// waves per EU: the most that leave the two-line kernel without spills up to m = 14
void dispatch_rphase_vpair(tdse_ctx* ctx, const Bucket& b, hipStream_t st) {
    switch (b.M) {
        case 2: launch_rphase_vpair<2, 4>(ctx, b, st); break;
        case 4: launch_rphase_vpair<4, 4>(ctx, b, st); break;
        case 6: launch_rphase_vpair<6, 4>(ctx, b, st); break;
        case 8: launch_rphase_vpair<8, 3>(ctx, b, st); break;
        case 10: launch_rphase_vpair<10, 2>(ctx, b, st); break;
        case 12: launch_rphase_vpair<12, 2>(ctx, b, st); break;
        case 14: launch_rphase_vpair<14, 2>(ctx, b, st); break;
        default: launch_rphase_vpair<16, 1>(ctx, b, st); break;
    }
}

// Two threads per chunk (each with half its rows) where the halves keep an even row count and
// the workgroup stays within 512 threads: half the registers, twice the waves.
template <bool REV>
void dispatch_vphase(tdse_ctx* ctx, const Bucket& b, int par, hipStream_t st) {
    const DevVec<int2>& pairs = b.vpairs[par];
    const bool split = (b.M == 8 || b.M == 12 || b.M == 16) && 2 * b.P <= 512;
    const dim3 grid((unsigned)pairs.count), block((unsigned)(split ? 2 * b.P : b.P));
    auto args = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, grid, block, 0, st, ctx->d_sh, pairs.ptr, (int)pairs.count, ctx->d_cl);
    };
    switch (b.M) {
        case 2: args(k_sh_vphase<2, REV, 1>); break;
        case 4: args(k_sh_vphase<4, REV, 1>); break;
        case 6: args(k_sh_vphase<6, REV, 1>); break;
        case 8: split ? args(k_sh_vphase<8, REV, 2>) : args(k_sh_vphase<8, REV, 1>); break;
        case 10: args(k_sh_vphase<10, REV, 1>); break;
        case 12: split ? args(k_sh_vphase<12, REV, 2>) : args(k_sh_vphase<12, REV, 1>); break;
        case 14: args(k_sh_vphase<14, REV, 1>); break;
        default: split ? args(k_sh_vphase<16, REV, 2>) : args(k_sh_vphase<16, REV, 1>); break;
    }
}

template <int M, bool IMAG, int METHOD>
void launch_line(tdse_ctx* ctx, const Bucket& b, int n_steps) {
    const size_t lds = (size_t)line_lds_rows(M, IMAG) * b.P * sizeof(double2) + 64 * sizeof(double);
    hipLaunchKernelGGL((k_line_steps<M, 512, IMAG, METHOD>), dim3((unsigned)b.lines.count), dim3(b.P), lds,
                       ctx->stream, ctx->d_line, b.lines.ptr, (int)b.lines.count, n_steps);
}

template <bool IMAG, int METHOD>
void dispatch_line_m(tdse_ctx* ctx, const Bucket& b, int n_steps) {
    switch (b.M) {
        case 2: launch_line<2, IMAG, METHOD>(ctx, b, n_steps); break;
        case 4: launch_line<4, IMAG, METHOD>(ctx, b, n_steps); break;
        case 6: launch_line<6, IMAG, METHOD>(ctx, b, n_steps); break;
        case 8: launch_line<8, IMAG, METHOD>(ctx, b, n_steps); break;
        case 10: launch_line<10, IMAG, METHOD>(ctx, b, n_steps); break;
        case 12: launch_line<12, IMAG, METHOD>(ctx, b, n_steps); break;
        case 14: launch_line<14, IMAG, METHOD>(ctx, b, n_steps); break;
        default: launch_line<16, IMAG, METHOD>(ctx, b, n_steps); break;
    }
}

// complex time steps take the general instance (ADI only: the host rejects the
// split operator with complex steps)
void dispatch_line(tdse_ctx* ctx, const Bucket& b, int n_steps) {
    if (b.imag) dispatch_line_m<false, LINE_ADI>(ctx, b, n_steps);
    else if (b.method == LINE_SIO_LEN) dispatch_line_m<true, LINE_SIO_LEN>(ctx, b, n_steps);
    else if (b.method == LINE_SIO_VEL) dispatch_line_m<true, LINE_SIO_VEL>(ctx, b, n_steps);
    else dispatch_line_m<true, LINE_ADI>(ctx, b, n_steps);
}

// The z-sweep solves two rho-lines per workgroup (2 Pz threads), explicit form
// (with two lines per workgroup the Cayley forms measured slower: re-reading the
// rows 0.44 ms, keeping them in LDS 0.64 ms per sweep against 0.40 ms).
template <int M, int MAXT, int WPE>
void launch_cyl_z(tdse_ctx* ctx, const Bucket& b) {
    const size_t lds = 2 * cn_lds_bytes(M, b.P, CN_EXPLICIT);
    hipLaunchKernelGGL((k_cyl_z<M, MAXT, WPE, CN_EXPLICIT>), dim3((unsigned)b.lines.count), dim3(2 * b.P), lds,
                       ctx->stream, ctx->d_cyl, b.lines.ptr, (int)b.lines.count);
}

void dispatch_cyl_z(tdse_ctx* ctx, const Bucket& b) {
    switch (b.M) {
        case 2: launch_cyl_z<2, 1024, 4>(ctx, b); break;
        case 4: launch_cyl_z<4, 1024, 4>(ctx, b); break;
        case 6: launch_cyl_z<6, 1024, 4>(ctx, b); break;
        case 8: launch_cyl_z<8, 1024, 4>(ctx, b); break;
        case 10: launch_cyl_z<10, 512, 3>(ctx, b); break;
        case 12: launch_cyl_z<12, 512, 2>(ctx, b); break;
        case 14: launch_cyl_z<14, 512, 2>(ctx, b); break;
        default: launch_cyl_z<16, 512, 2>(ctx, b); break;
    }
}

template <int M, int LANES, int MAXT, int WPE>
void launch_cyl_rho(tdse_ctx* ctx, const Bucket& b, int mode) {
    const size_t lds = 3 * (size_t)b.P * LANES * sizeof(double2) + 16 * sizeof(double);
    const dim3 grid((unsigned)b.lines.count), block((unsigned)(b.P * LANES));
    if (mode == CYL_PRE)
        hipLaunchKernelGGL((k_cyl_rho<M, LANES, CYL_PRE, MAXT, WPE>), grid, block, lds, ctx->stream, ctx->d_cyl,
                           b.lines.ptr, (int)b.lines.count);
    else
        hipLaunchKernelGGL((k_cyl_rho<M, LANES, CYL_MAIN, MAXT, WPE>), grid, block, lds, ctx->stream, ctx->d_cyl,
                           b.lines.ptr, (int)b.lines.count);
}

// rho-sweep shapes (rows, lines per workgroup, block limit, minimum waves per
// EU); the layout chooser of the host package picks one of these.
void dispatch_cyl_rho(tdse_ctx* ctx, const Bucket& b, int mode) {
    if (b.M == 4 && b.lanes == 4) launch_cyl_rho<4, 4, 1024, 4>(ctx, b, mode);
    else if (b.M == 8 && b.lanes == 4) launch_cyl_rho<8, 4, 512, 4>(ctx, b, mode);
    else if (b.M == 4 && b.lanes == 8) launch_cyl_rho<4, 8, 1024, 4>(ctx, b, mode);
    else if (b.M == 8 && b.lanes == 8) launch_cyl_rho<8, 8, 1024, 4>(ctx, b, mode);
    else launch_cyl_rho<16, 4, 512, 2>(ctx, b, mode);
}

// Launch one full step of every simulation kind onto the context stream.
// cyl_pre: include the cylindrical pre-sweep (needed while any cylindrical
// simulation may start a step from an exact state).
int launch_step(tdse_ctx* ctx, bool cyl_pre) {
    const int nw = (int)ctx->waves.count;
    if (nw > 0) {
        prof_begin(ctx, 0);
        // the instance that reduces the norm in every l needs more registers: only when asked for
        const int nl = ctx->n_len_waves, nv = nw - nl;
        if (nl > 0) {
            auto lphase = ctx->sh_norm_by_l ? k_sh_lphase<true, false> : k_sh_lphase<false, false>;
            hipLaunchKernelGGL(lphase, dim3((nl * TDSE_WAVE + 255) / 256), dim3(256), 0, ctx->stream, ctx->d_sh,
                               ctx->waves.ptr, nl, ctx->d_cl, 0);
        }
        if (nv > 0) {
            auto lphase = ctx->sh_norm_by_l ? k_sh_lphase<true, true> : k_sh_lphase<false, true>;
            hipLaunchKernelGGL(lphase, dim3((nv * TDSE_WAVE + 255) / 256), dim3(256), 0, ctx->stream, ctx->d_sh,
                               ctx->waves.ptr + nl, nv, ctx->d_clv, 0);
        }
        prof_end(ctx);
        // The observables read only what the L-phase wrote (per-wave partials, snapshot rows), which
        // nothing else touches before the next replay: they can run beside the R-phase.
        const int nj = (int)ctx->jobs.count;
        auto observe = [&](hipStream_t st) {
            hipLaunchKernelGGL(k_sh_observe, dim3(nj), dim3(256), 0, st, ctx->d_sh, ctx->jobs.ptr, nj);
        };
        // One bucket's chain: velocity-gauge h2 factors of its even, then odd l-pairs
        // (before the CN solve), the R-phase, the factors in reverse (after it). The
        // chains of different buckets touch different simulations.
        auto chain = [&](const Bucket& b, hipStream_t st) {
            for (int par = 0; par < 2; ++par)
                if (b.vpairs[par].count) {
                    prof_begin(ctx, 4);
                    dispatch_vphase<false>(ctx, b, par, st);
                    prof_end(ctx);
                }
            if (b.lines.count) {
                prof_begin(ctx, 1);
                dispatch_rphase(ctx, b, st);
                prof_end(ctx);
            }
            if (b.fpairs.count) {
                prof_begin(ctx, 5);
                dispatch_rphase_vpair(ctx, b, st);
                prof_end(ctx);
            }
            for (int par = 1; par >= 0; --par)
                if (b.vpairs[par].count) {
                    prof_begin(ctx, 4);
                    dispatch_vphase<true>(ctx, b, par, st);
                    prof_end(ctx);
                }
        };
        if (ctx->profiling || ctx->sh_buckets.size() < 2) {
            observe(ctx->stream);
            for (auto& b : ctx->sh_buckets) chain(b, ctx->stream);
        } else {
            // independent buckets run concurrently (largest first, round robin
            // over the context stream and the side streams) so that small
            // buckets fill the tails of large ones; fork and join by events
            HIP_TRY(ctx, hipEventRecord(ctx->ev_fork, ctx->stream));
            for (int j = 0; j < kSideStreams; ++j) HIP_TRY(ctx, hipStreamWaitEvent(ctx->side[j], ctx->ev_fork, 0));
            // the observables on the last side stream, which the round robin leaves the least work
            observe(ctx->side[kSideStreams - 1]);
            for (size_t i = 0; i < ctx->sh_order.size(); ++i) {
                const int j = (int)(i % (kSideStreams + 1));
                chain(ctx->sh_buckets[ctx->sh_order[i]], j == 0 ? ctx->stream : ctx->side[j - 1]);
            }
            for (int j = 0; j < kSideStreams; ++j) {
                HIP_TRY(ctx, hipEventRecord(ctx->ev_join[j], ctx->side[j]));
                HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_join[j], 0));
            }
        }
        // AlternatingDirectionImplicit simulations: explicit H0, the l-direction factor, implicit H0.
        // Within each of the three sweeps the buckets are independent: they run concurrently
        // over the context stream and the side streams (fork and join by events), and each
        // sweep starts after the whole previous one.
        auto fan = [&](std::vector<Bucket>& bs, auto&& launch) -> int {
            if (ctx->profiling || bs.size() < 2) {
                for (auto& b : bs) launch(b, ctx->stream);
                return TDSE_OK;
            }
            HIP_TRY(ctx, hipEventRecord(ctx->ev_fork, ctx->stream));
            for (int j = 0; j < kSideStreams; ++j) HIP_TRY(ctx, hipStreamWaitEvent(ctx->side[j], ctx->ev_fork, 0));
            for (size_t i = 0; i < bs.size(); ++i) {
                const int j = (int)(i % (kSideStreams + 1));
                launch(bs[i], j == 0 ? ctx->stream : ctx->side[j - 1]);
            }
            for (int j = 0; j < kSideStreams; ++j) {
                HIP_TRY(ctx, hipEventRecord(ctx->ev_join[j], ctx->side[j]));
                HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_join[j], 0));
            }
            return TDSE_OK;
        };
        if (!ctx->sh_adi_buckets.empty()) {
            int rc;
            if ((rc = fan(ctx->sh_adi_buckets, [&](const Bucket& b, hipStream_t st) { dispatch_adi_r<true>(ctx, b, st); })))
                return rc;
            if ((rc = fan(ctx->sh_adi_lbuckets, [&](const Bucket& b, hipStream_t st) { dispatch_adi_lp(ctx, b, st); })))
                return rc;
            const int na = (int)ctx->adi_waves.count;
            if (na > 0)
                hipLaunchKernelGGL(k_sh_adi_l, dim3((na * TDSE_WAVE + 255) / 256), dim3(256), 0, ctx->stream,
                                   ctx->d_sh, ctx->adi_waves.ptr, na, ctx->d_cl);
            if ((rc = fan(ctx->sh_adi_buckets, [&](const Bucket& b, hipStream_t st) { dispatch_adi_r<false>(ctx, b, st); })))
                return rc;
        }
        const int ns = (int)ctx->sh.size();
        hipLaunchKernelGGL(k_sh_advance, dim3((ns + 255) / 256), dim3(256), 0, ctx->stream, ctx->d_sh, ns, ctx->d_ctl);
    }
    for (auto& b : ctx->line_buckets) dispatch_line(ctx, b, 1);
    if (!ctx->cyl_r_buckets.empty()) {
        const int nj = (int)ctx->cyl_jobs.count;
        if (cyl_pre) {
            for (auto& b : ctx->cyl_r_buckets) dispatch_cyl_rho(ctx, b, CYL_PRE);
            hipLaunchKernelGGL(k_cyl_observe, dim3(nj), dim3(256), 0, ctx->stream, ctx->d_cyl, ctx->cyl_jobs.ptr, nj,
                               0);
        }
        for (auto& b : ctx->cyl_z_buckets) {
            prof_begin(ctx, 2);
            dispatch_cyl_z(ctx, b);
            prof_end(ctx);
        }
        for (auto& b : ctx->cyl_r_buckets) {
            prof_begin(ctx, 3);
            dispatch_cyl_rho(ctx, b, CYL_MAIN);
            prof_end(ctx);
        }
        hipLaunchKernelGGL(k_cyl_observe, dim3(nj), dim3(256), 0, ctx->stream, ctx->d_cyl, ctx->cyl_jobs.ptr, nj, 1);
        const int nc = (int)ctx->cyl.size();
        hipLaunchKernelGGL(k_cyl_advance, dim3((nc + 255) / 256), dim3(256), 0, ctx->stream, ctx->d_cyl, nc);
    }
    HIP_TRY(ctx, hipGetLastError());
    return TDSE_OK;
}

// Capture `steps` consecutive steps into one graph.
int capture_steps(tdse_ctx* ctx, bool cyl_pre, int steps, hipGraph_t* graph, hipGraphExec_t* exec) {
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int i = 0; i < steps && !rc; ++i) rc = launch_step(ctx, cyl_pre);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(ctx, TDSE_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    *graph = g;
    HIP_TRY(ctx, hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return TDSE_OK;
}

// The one-step graph and the kGraphSteps-step graph of the current tables.
int ensure_graph(tdse_ctx* ctx, bool cyl_pre) {
    if (ctx->exec && (kGraphSteps <= 1 || ctx->exec_k) && ctx->graph_cyl_pre == cyl_pre) return TDSE_OK;
    drop_graph(ctx);
    ctx->graph_cyl_pre = cyl_pre;
    int rc = capture_steps(ctx, cyl_pre, 1, &ctx->graph, &ctx->exec);
    if (!rc && kGraphSteps > 1) rc = capture_steps(ctx, cyl_pre, kGraphSteps, &ctx->graph_k, &ctx->exec_k);
    return rc;
}

template <class T>
int copy_to_device(tdse_ctx* ctx, std::vector<void*>& blocks, const T* host, size_t count, T** out) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    void* p = ctx->arena.alloc(bytes);
    if (!p) return fail(ctx, TDSE_ERR_NOMEM, "device pool exhausted");
    blocks.push_back(p);
    if (host && count) HIP_TRY(ctx, hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    else HIP_TRY(ctx, hipMemsetAsync(p, 0, bytes, ctx->stream));
    *out = static_cast<T*>(p);
    return TDSE_OK;
}

// Plain streaming copy: one complex128 read and one write per element, the
// grid covering the buffer once (one element per thread up to 2^28 elements).
// On MI355X this form measured 6.2 TB/s for 1 GiB each way; a grid-stride
// loop over 8192 workgroups reached 4.5 TB/s.
__global__ void __launch_bounds__(256) k_stream_copy(const double2* __restrict__ src, double2* __restrict__ dst,
                                                     size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) st2(dst, i, ld2(src, i));
}

// Scratch of the allocation-free Thomas entry point: the primed super-diagonal
// of up to kNoAllocPoints unknowns, then kNoAllocSystems int64 offsets and as
// many int64 lengths (33 MiB in all). Each device holds its own copy until the
// process ends. In host code the name `scratch` is only a host-side stand-in:
// kernels and copies use the device address that hipGetSymbolAddress returns
// for the current device (device_scratch), never the name itself.
constexpr int64_t kNoAllocPoints = int64_t(1) << 21;
constexpr int64_t kNoAllocSystems = int64_t(1) << 16;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double2 scratch[kNoAllocPoints + kNoAllocSystems];

std::mutex g_scratch_mu;
double2* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current), looked up
// once per device. Returns nullptr with *err set on failure.
double2* device_scratch(int device, hipError_t* err) {
    *err = hipSuccess;
    if (device < 0 || device >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        double2* p = nullptr;
        *err = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
        if (*err != hipSuccess) return nullptr;
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

void free_blocks(tdse_ctx* ctx, std::vector<void*>& blocks) {
    for (void* p : blocks) ctx->arena.dealloc(p);
    blocks.clear();
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

int tdse_abi_version(void) { return kAbiVersion; }

const char* tdse_last_error(const tdse_ctx* ctx) { return ctx ? ctx->err.c_str() : g_lib_error.c_str(); }

int tdse_ctx_create(int device, size_t pool_bytes, tdse_ctx** out) {
    if (!out) return fail(nullptr, TDSE_ERR_ARG, "out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(nullptr, TDSE_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= count) return fail(nullptr, TDSE_ERR_ARG, "device index out of range");
    if (hipSetDevice(device) != hipSuccess) return fail(nullptr, TDSE_ERR_HIP, "hipSetDevice failed");
    tdse_ctx* ctx = new tdse_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return fail(nullptr, TDSE_ERR_HIP, "hipStreamCreate failed");
    }
    if (!ctx->arena.init(pool_bytes)) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return fail(nullptr, TDSE_ERR_NOMEM, "could not reserve the device memory pool");
    }
    std::vector<double> cl(kMaxL), clv(kMaxL);
    for (int l = 0; l < kMaxL; ++l) {
        cl[l] = (l + 1.0) / std::sqrt((2.0 * l + 1.0) * (2.0 * l + 3.0));
        clv[l] = cl[l] * (l + 1.0);
    }
    bool ok = hipMalloc(&ctx->d_cl, kMaxL * sizeof(double)) == hipSuccess &&
              hipMemcpy(ctx->d_cl, cl.data(), kMaxL * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
              hipMalloc(&ctx->d_clv, kMaxL * sizeof(double)) == hipSuccess &&
              hipMemcpy(ctx->d_clv, clv.data(), kMaxL * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
              hipMalloc(&ctx->d_sh, kSlotCapacity * sizeof(ShSim)) == hipSuccess &&
              hipMalloc(&ctx->d_ctl, sizeof(int)) == hipSuccess && hipMemset(ctx->d_ctl, 0, sizeof(int)) == hipSuccess &&
              hipMalloc(&ctx->d_line, kSlotCapacity * sizeof(LineSim)) == hipSuccess &&
              hipMalloc(&ctx->d_cyl, kSlotCapacity * sizeof(CylSim)) == hipSuccess &&
              hipMemset(ctx->d_sh, 0, kSlotCapacity * sizeof(ShSim)) == hipSuccess &&
              hipMemset(ctx->d_line, 0, kSlotCapacity * sizeof(LineSim)) == hipSuccess &&
              hipMemset(ctx->d_cyl, 0, kSlotCapacity * sizeof(CylSim)) == hipSuccess &&
              hipEventCreate(&ctx->ev_start) == hipSuccess && hipEventCreate(&ctx->ev_stop) == hipSuccess &&
              hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) == hipSuccess;
    for (int j = 0; ok && j < kSideStreams; ++j)
        ok = hipStreamCreateWithFlags(&ctx->side[j], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&ctx->ev_join[j], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        tdse_ctx_destroy(ctx);
        return fail(nullptr, TDSE_ERR_HIP, "context allocation failed");
    }
    hipError_t se = hipSuccess;
    ctx->scratch = device_scratch(device, &se);
    if (!ctx->scratch) {
        tdse_ctx_destroy(ctx);
        return fail(nullptr, TDSE_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se));
    }
    *out = ctx;
    return TDSE_OK;
}

int tdse_ctx_destroy(tdse_ctx* ctx) {
    if (!ctx) return TDSE_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    drop_graph(ctx);
    for (auto* v : {&ctx->sh_buckets, &ctx->sh_adi_buckets, &ctx->sh_adi_lbuckets, &ctx->line_buckets,
                    &ctx->cyl_z_buckets, &ctx->cyl_r_buckets})
        for (auto& b : *v) free_bucket(b);
    if (ctx->adi_waves.ptr) (void)hipFree(ctx->adi_waves.ptr);
    if (ctx->cyl_jobs.ptr) (void)hipFree(ctx->cyl_jobs.ptr);
    if (ctx->d_cyl) (void)hipFree(ctx->d_cyl);
    if (ctx->waves.ptr) (void)hipFree(ctx->waves.ptr);
    if (ctx->jobs.ptr) (void)hipFree(ctx->jobs.ptr);
    if (ctx->d_cl) (void)hipFree(ctx->d_cl);
    if (ctx->d_clv) (void)hipFree(ctx->d_clv);
    if (ctx->d_sh) (void)hipFree(ctx->d_sh);
    if (ctx->d_ctl) (void)hipFree(ctx->d_ctl);
    if (ctx->d_line) (void)hipFree(ctx->d_line);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    for (int j = 0; j < kSideStreams; ++j) {
        if (ctx->ev_join[j]) (void)hipEventDestroy(ctx->ev_join[j]);
        if (ctx->side[j]) (void)hipStreamDestroy(ctx->side[j]);
    }
    ctx->arena.release();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return TDSE_OK;
}

int tdse_sync(tdse_ctx* ctx) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return TDSE_OK;
}

int tdse_pool_usage(tdse_ctx* ctx, size_t* used_bytes, size_t* total_bytes) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (used_bytes) *used_bytes = ctx->arena.used;
    if (total_bytes) *total_bytes = ctx->arena.size;
    return TDSE_OK;
}

int tdse_copy_probe(tdse_ctx* ctx, size_t bytes, int launches, double* avg_ms) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (bytes == 0 || bytes % sizeof(double2) != 0 || launches < 1)
        return fail(ctx, TDSE_ERR_ARG, "tdse_copy_probe: bytes must be a positive multiple of 16, launches >= 1");
    std::vector<void*> blocks;
    double2 *src, *dst;
    const size_t n = bytes / sizeof(double2);
    int rc = copy_to_device<double2>(ctx, blocks, nullptr, n, &src);
    if (rc == TDSE_OK) rc = copy_to_device<double2>(ctx, blocks, nullptr, n, &dst);
    if (rc != TDSE_OK) {
        free_blocks(ctx, blocks);
        return rc;
    }
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, (size_t)1 << 20);
    hipError_t e = hipEventRecord(ctx->ev_start, ctx->stream);
    for (int i = 0; i < launches && e == hipSuccess; ++i) {
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, ctx->stream, src, dst, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev_stop, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev_stop);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
    free_blocks(ctx, blocks);
    if (e != hipSuccess) return fail(ctx, TDSE_ERR_HIP, std::string("tdse_copy_probe: ") + hipGetErrorString(e));
    if (avg_ms) *avg_ms = ms / launches;
    return TDSE_OK;
}

int tdse_tdma_batched(tdse_ctx* ctx, const double* sub, const double* diag, const double* sup, const double* d,
                      double* x, const int64_t* offsets, const int64_t* lengths, int64_t n_systems) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (!sub || !diag || !sup || !d || !x || !offsets || !lengths || n_systems < 0)
        return fail(ctx, TDSE_ERR_ARG, "tdse_tdma_batched: NULL array or negative count");
    if (n_systems == 0) return TDSE_OK;
    int64_t total = 0;
    for (int64_t i = 0; i < n_systems; ++i) {
        if (lengths[i] < 1 || offsets[i] < 0) return fail(ctx, TDSE_ERR_ARG, "tdse_tdma_batched: bad system extent");
        total = std::max(total, offsets[i] + lengths[i]);
    }
    std::vector<void*> blocks;
    double2 *da, *db, *dc, *dd, *dx, *dcp;
    int64_t *doff, *dlen;
    int rc = 0;
    const size_t n = (size_t)total;
    if ((rc = copy_to_device(ctx, blocks, reinterpret_cast<const double2*>(sub), n, &da)) ||
        (rc = copy_to_device(ctx, blocks, reinterpret_cast<const double2*>(diag), n, &db)) ||
        (rc = copy_to_device(ctx, blocks, reinterpret_cast<const double2*>(sup), n, &dc)) ||
        (rc = copy_to_device(ctx, blocks, reinterpret_cast<const double2*>(d), n, &dd)) ||
        (rc = copy_to_device<double2>(ctx, blocks, nullptr, n, &dx)) ||
        (rc = copy_to_device<double2>(ctx, blocks, nullptr, n, &dcp)) ||
        (rc = copy_to_device(ctx, blocks, offsets, (size_t)n_systems, &doff)) ||
        (rc = copy_to_device(ctx, blocks, lengths, (size_t)n_systems, &dlen))) {
        (void)hipStreamSynchronize(ctx->stream);
        free_blocks(ctx, blocks);
        return rc;
    }
    const int threads = 64;
    const int64_t grid = (n_systems + threads - 1) / threads;
    hipLaunchKernelGGL(k_tdma_lines, dim3((unsigned)grid), dim3(threads), 0, ctx->stream, da, db, dc, dd, dx, dcp, doff,
                       dlen, n_systems);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(x, dx, n * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    free_blocks(ctx, blocks);
    if (e != hipSuccess) return fail(ctx, TDSE_ERR_HIP, std::string("tdse_tdma_batched: ") + hipGetErrorString(e));
    return TDSE_OK;
}

int tdse_tdma(tdse_ctx* ctx, const double* sub, const double* diag, const double* sup, const double* d, double* x,
              int64_t n) {
    if (n < 1) return fail(ctx, TDSE_ERR_ARG, "tdse_tdma: n must be >= 1");
    const int64_t off = 0;
    return tdse_tdma_batched(ctx, sub, diag, sup, d, x, &off, &n, 1);
}

int tdse_tdma_batched_noalloc(tdse_ctx* ctx, const double* sub, const double* diag, const double* sup,
                              const double* d, double* x, const int64_t* offsets, const int64_t* lengths,
                              int64_t n_systems, int64_t total) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (!sub || !diag || !sup || !d || !x || !offsets || !lengths || n_systems < 0 || total < 0)
        return fail(ctx, TDSE_ERR_ARG, "tdse_tdma_batched_noalloc: NULL array or negative count");
    if (total > kNoAllocPoints)
        return fail(ctx, TDSE_ERR_ARG, "tdse_tdma_batched_noalloc: total " + std::to_string(total) +
                                           " exceeds the scratch maximum of " + std::to_string(kNoAllocPoints) +
                                           " unknowns");
    if (n_systems > kNoAllocSystems)
        return fail(ctx, TDSE_ERR_ARG, "tdse_tdma_batched_noalloc: " + std::to_string(n_systems) +
                                           " systems exceed the scratch maximum of " +
                                           std::to_string(kNoAllocSystems));
    for (int64_t i = 0; i < n_systems; ++i)
        if (lengths[i] < 1 || offsets[i] < 0 || offsets[i] > total - lengths[i])
            return fail(ctx, TDSE_ERR_ARG, "tdse_tdma_batched_noalloc: system " + std::to_string(i) +
                                               " lies outside [0, total)");
    if (n_systems == 0) return TDSE_OK;
    if (!ctx->scratch) return fail(ctx, TDSE_ERR_STATE, "tdse_tdma_batched_noalloc: no device scratch");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    double2* cp = ctx->scratch;
    int64_t* doff = reinterpret_cast<int64_t*>(cp + kNoAllocPoints);
    int64_t* dlen = doff + kNoAllocSystems;
    const size_t nb = (size_t)n_systems * sizeof(int64_t);
    HIP_TRY(ctx, hipMemcpyAsync(doff, offsets, nb, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(dlen, lengths, nb, hipMemcpyHostToDevice, ctx->stream));
    const int threads = 64;
    const int64_t grid = (n_systems + threads - 1) / threads;
    hipLaunchKernelGGL(k_tdma_lines, dim3((unsigned)grid), dim3(threads), 0, ctx->stream,
                       reinterpret_cast<const double2*>(sub), reinterpret_cast<const double2*>(diag),
                       reinterpret_cast<const double2*>(sup), reinterpret_cast<const double2*>(d),
                       reinterpret_cast<double2*>(x), cp, doff, dlen, n_systems);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return TDSE_OK;
}

int tdse_sh_add(tdse_ctx* ctx, const tdse_sh_desc* desc, const double* tables, const double* fs, const double* ts,
                const double* g0, const double* basis, const int32_t* basis_l, int32_t* slot) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (!desc || !tables || !fs || !ts || !g0 || !slot) return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: NULL argument");
    const tdse_sh_desc& D = *desc;
    const int P = D.chunks, m = D.rows;
    if (D.r_points < 2 || D.l_bound < 1 || D.l_bound > kMaxL) return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: bad mesh size");
    if (P < TDSE_WAVE || P % TDSE_WAVE) return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: chunks must be a positive multiple of 64");
    if (m < 2 || m > 16 || (m & 1)) return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: rows must be even and in [2, 16]");
    if (P > 512) return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: at most 512 chunks");
    if ((int64_t)P * m < D.r_points) return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: chunks * rows < r_points");
    if (D.n_steps < 1) return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: n_steps must be >= 1");
    if (D.store_every == 0 || D.store_every < -1) return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: store_every must be >= 1 or -1");
    if (D.n_basis < 0 || (D.n_basis > 0 && (!basis || !basis_l))) return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: basis missing");
    if (D.lb < 1 || D.lb > D.l_bound) return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: lb out of range");
    if (D.gauge != 0 && D.gauge != 1) return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: gauge must be 0 (length) or 1 (velocity)");
    if (D.gauge == 1 && (D.r_points < 3 || D.l_bound < 2))
        return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: the velocity gauge needs r_points >= 3 and l_bound >= 2");
    if (D.method != 0 && D.method != 1) return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: method must be 0 or 1");
    if (D.method == 1 && D.gauge != 0)
        return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: AlternatingDirectionImplicit runs the length gauge");
    for (int b = 0; b < D.n_basis; ++b)
        if (basis_l[b] < 0 || basis_l[b] >= D.lb) return fail(ctx, TDSE_ERR_ARG, "tdse_sh_add: basis l must be < lb");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int s = -1;
    for (int i = 0; i < (int)ctx->sh.size(); ++i)
        if (!ctx->sh[i].live) {
            s = i;
            break;
        }
    if (s < 0) {
        if ((int)ctx->sh.size() >= kSlotCapacity) return fail(ctx, TDSE_ERR_NOMEM, "no free simulation slot");
        ctx->sh.emplace_back();
        s = (int)ctx->sh.size() - 1;
    }
    ShHost& H = ctx->sh[s];
    H.blocks.clear();
    H.vel = D.gauge == 1;
    H.adi = D.method == 1;
    ShSim& S = H.dev;
    S = ShSim{};
    S.n = D.r_points;
    S.L = D.l_bound;
    S.P = P;
    S.m = m;
    S.npad = P * m;
    S.N = D.n_steps;
    S.store_every = D.store_every;
    S.LB = D.lb;
    S.n_basis = D.n_basis;
    S.nwaves = (S.npad / TDSE_WAVE) * sh_segments(S.L);
    S.norm_by_l = D.norm_by_l ? 1 : 0;
    S.corr0 = D.corr0;
    S.delta_r = D.delta_r;
    S.cu = D.centrifugal;
    S.step = 0;
    S.active = 1;
    H.n_data = n_data_points(S.N, S.store_every);
    const size_t npad = S.npad;
    double2 *g, *g2, *snap, *oip, *ta;
    double *tab, *dfs, *dts, *part, *on, *oln, *bas = nullptr, *pl = nullptr, *onl = nullptr, *dfa = nullptr;
    int* bl = nullptr;
    int rc = 0;
    std::vector<double2> ta_host(npad);  // (T_j, A_j) interleaved for the R-phase
    for (size_t p = 0; p < npad; ++p) ta_host[p] = make_double2(tables[SH_T * npad + p], tables[SH_AN * npad + p]);
    if ((rc = copy_to_device(ctx, H.blocks, reinterpret_cast<const double2*>(g0), (size_t)S.L * npad, &g)) ||
        (rc = copy_to_device(ctx, H.blocks, ta_host.data(), npad, &ta)) ||
        (rc = copy_to_device<double2>(ctx, H.blocks, nullptr, (size_t)S.L * npad, &g2)) ||
        (rc = copy_to_device(ctx, H.blocks, tables, (SH_NTAB + (H.vel ? 1 : 0)) * npad, &tab)) ||
        // ADI: the L-phase sees a zero field (it only masks and observes); the field goes to fa
        (rc = copy_to_device(ctx, H.blocks, H.adi ? nullptr : fs, (size_t)S.N, &dfs)) ||
        (H.adi && (rc = copy_to_device(ctx, H.blocks, fs, (size_t)S.N, &dfa))) ||
        (rc = copy_to_device(ctx, H.blocks, ts, (size_t)S.N, &dts)) ||
        (rc = copy_to_device<double2>(ctx, H.blocks, nullptr, (size_t)S.LB * npad, &snap)) ||
        (rc = copy_to_device<double>(ctx, H.blocks, nullptr, 2 * (size_t)S.nwaves, &part)) ||
        (rc = copy_to_device<double>(ctx, H.blocks, nullptr, (size_t)H.n_data, &on)) ||
        (rc = copy_to_device<double>(ctx, H.blocks, nullptr, (size_t)H.n_data, &oln)) ||
        (rc = copy_to_device<double2>(ctx, H.blocks, nullptr, (size_t)H.n_data * std::max(1, S.n_basis), &oip)) ||
        (S.n_basis > 0 && ((rc = copy_to_device(ctx, H.blocks, basis, (size_t)S.n_basis * npad, &bas)) ||
                           (rc = copy_to_device(ctx, H.blocks, basis_l, (size_t)S.n_basis, &bl)))) ||
        (S.norm_by_l &&
         ((rc = copy_to_device<double>(ctx, H.blocks, nullptr, (size_t)S.nwaves * kShSegRows, &pl)) ||
          (rc = copy_to_device<double>(ctx, H.blocks, nullptr, (size_t)H.n_data * S.L, &onl))))) {
        (void)hipStreamSynchronize(ctx->stream);
        free_blocks(ctx, H.blocks);
        return rc;
    }
    S.g = g;
    S.g2 = g2;
    S.ta = ta;
    S.tab = tab;
    S.fs = dfs;
    S.ts = dts;
    S.snap = snap;
    S.part = part;
    S.out_norm = on;
    S.out_lnorm = oln;
    S.out_ip = oip;
    S.basis = bas;
    S.basis_l = bl;
    S.part_l = pl;
    S.out_nl = onl;
    S.fa = dfa;
    S.adi = H.adi ? 1 : 0;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->d_sh + s, &S, sizeof(ShSim), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // host buffers may be released after return
    H.live = true;
    ctx->dirty = true;
    *slot = s;
    return TDSE_OK;
}

int tdse_line_add(tdse_ctx* ctx, const tdse_line_desc* desc, const double* tables, const double* wt, const double* fs,
                  const double* g0, const double* basis, int32_t* slot) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (!desc || !tables || !wt || !fs || !g0 || !slot) return fail(ctx, TDSE_ERR_ARG, "tdse_line_add: NULL argument");
    const tdse_line_desc& D = *desc;
    const int P = D.chunks, m = D.rows;
    if (D.z_points < 2) return fail(ctx, TDSE_ERR_ARG, "tdse_line_add: z_points must be >= 2");
    if (P < TDSE_WAVE || P % TDSE_WAVE) return fail(ctx, TDSE_ERR_ARG, "tdse_line_add: chunks must be a positive multiple of 64");
    if (m < 2 || m > 16 || (m & 1)) return fail(ctx, TDSE_ERR_ARG, "tdse_line_add: rows must be even and in [2, 16]");
    if (P > 512) return fail(ctx, TDSE_ERR_ARG, "tdse_line_add: chunks must be <= 512");
    if ((int64_t)P * m < D.z_points) return fail(ctx, TDSE_ERR_ARG, "tdse_line_add: chunks * rows < z_points");
    if (D.n_steps < 1) return fail(ctx, TDSE_ERR_ARG, "tdse_line_add: n_steps must be >= 1");
    if (D.store_every == 0 || D.store_every < -1) return fail(ctx, TDSE_ERR_ARG, "tdse_line_add: store_every must be >= 1 or -1");
    if (D.n_basis < 0 || (D.n_basis > 0 && !basis)) return fail(ctx, TDSE_ERR_ARG, "tdse_line_add: basis missing");
    if (D.method < 0 || D.method > 2) return fail(ctx, TDSE_ERR_ARG, "tdse_line_add: method must be 0, 1 or 2");
    if (D.method != 0 && D.renormalize)
        return fail(ctx, TDSE_ERR_ARG, "tdse_line_add: the split operator does not renormalise");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int s = -1;
    for (int i = 0; i < (int)ctx->line.size(); ++i)
        if (!ctx->line[i].live) {
            s = i;
            break;
        }
    if (s < 0) {
        if ((int)ctx->line.size() >= kSlotCapacity) return fail(ctx, TDSE_ERR_NOMEM, "no free simulation slot");
        ctx->line.emplace_back();
        s = (int)ctx->line.size() - 1;
    }
    LineHost& H = ctx->line[s];
    H.blocks.clear();
    H.complex_steps = false;
    for (int i = 0; i < D.n_steps; ++i)
        if (wt[2 * i] != 0.0) H.complex_steps = true;
    if (D.method != 0 && H.complex_steps)
        return fail(ctx, TDSE_ERR_ARG, "tdse_line_add: the split operator needs real time steps");
    LineSim& S = H.dev;
    S = LineSim{};
    S.n = D.z_points;
    S.P = P;
    S.m = m;
    S.npad = P * m;
    S.N = D.n_steps;
    S.store_every = D.store_every;
    S.n_basis = D.n_basis;
    S.renormalize = D.renormalize;
    S.delta_z = D.delta_z;
    S.off = D.offdiag;
    S.zq0 = D.zq_first;
    S.dzq = D.zq_step;
    S.vq = D.vel_factor;
    S.method = D.method;
    S.step = 0;
    S.active = 1;
    H.n_data = n_data_points(S.N, S.store_every);
    const size_t npad = S.npad;
    double2 *g, *dwt, *oip, *bas = nullptr;
    double *tab, *dfs, *on;
    int rc = 0;
    if ((rc = copy_to_device(ctx, H.blocks, reinterpret_cast<const double2*>(g0), npad, &g)) ||
        (rc = copy_to_device(ctx, H.blocks, tables, LN_NTAB * npad, &tab)) ||
        (rc = copy_to_device(ctx, H.blocks, reinterpret_cast<const double2*>(wt), (size_t)S.N, &dwt)) ||
        (rc = copy_to_device(ctx, H.blocks, fs, (size_t)S.N, &dfs)) ||
        (rc = copy_to_device<double>(ctx, H.blocks, nullptr, (size_t)H.n_data, &on)) ||
        (rc = copy_to_device<double2>(ctx, H.blocks, nullptr, (size_t)H.n_data * std::max(1, S.n_basis), &oip)) ||
        (S.n_basis > 0 &&
         (rc = copy_to_device(ctx, H.blocks, reinterpret_cast<const double2*>(basis), (size_t)S.n_basis * npad, &bas)))) {
        (void)hipStreamSynchronize(ctx->stream);
        free_blocks(ctx, H.blocks);
        return rc;
    }
    S.g = g;
    S.tab = tab;
    S.wt = dwt;
    S.fs = dfs;
    S.out_norm = on;
    S.out_ip = oip;
    S.basis = bas;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->d_line + s, &S, sizeof(LineSim), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // host buffers may be released after return
    H.live = true;
    ctx->dirty = true;
    *slot = s;
    return TDSE_OK;
}

int tdse_cyl_add(tdse_ctx* ctx, const tdse_cyl_desc* desc, const double* zt, const double* rt, const double* roff,
                 const double* fs, const double* ts, const double* g0, const double* basis, int32_t* slot) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (!desc || !zt || !rt || !roff || !fs || !ts || !g0 || !slot)
        return fail(ctx, TDSE_ERR_ARG, "tdse_cyl_add: NULL argument");
    const tdse_cyl_desc& D = *desc;
    if (D.z_points < 2 || D.rho_points < 2) return fail(ctx, TDSE_ERR_ARG, "tdse_cyl_add: bad mesh size");
    const int Pz = D.z_chunks, Mz = D.z_rows, Cr = D.rho_chunks, Mr = D.rho_rows, lanes = D.lanes;
    if (Pz < TDSE_WAVE || Pz % TDSE_WAVE) return fail(ctx, TDSE_ERR_ARG, "tdse_cyl_add: z_chunks must be a positive multiple of 64");
    if (Mz < 2 || Mz > 16 || (Mz & 1)) return fail(ctx, TDSE_ERR_ARG, "tdse_cyl_add: z_rows must be even and in [2, 16]");
    if ((Mz <= 8 && Pz > 512) || (Mz > 8 && Pz > 256)) return fail(ctx, TDSE_ERR_ARG, "tdse_cyl_add: too many z chunks for z rows");
    if ((int64_t)Pz * Mz < D.z_points) return fail(ctx, TDSE_ERR_ARG, "tdse_cyl_add: z_chunks * z_rows < z_points");
    const bool shape_ok = (Mr == 4 && lanes == 4 && Cr * 4 <= 1024) || (Mr == 8 && lanes == 4 && Cr * 4 <= 512) ||
                          (Mr == 4 && lanes == 8 && Cr * 8 <= 1024) || (Mr == 8 && lanes == 8 && Cr * 8 <= 1024) ||
                          (Mr == 16 && lanes == 4 && Cr * 4 <= 512);
    if (!shape_ok || Cr < 1 || (Cr * lanes) % TDSE_WAVE)
        return fail(ctx, TDSE_ERR_ARG, "tdse_cyl_add: unsupported rho layout (rows, lanes, chunks)");
    if ((int64_t)Cr * Mr < D.rho_points) return fail(ctx, TDSE_ERR_ARG, "tdse_cyl_add: rho_chunks * rho_rows < rho_points");
    if (D.n_steps < 1) return fail(ctx, TDSE_ERR_ARG, "tdse_cyl_add: n_steps must be >= 1");
    if (D.store_every == 0 || D.store_every < -1) return fail(ctx, TDSE_ERR_ARG, "tdse_cyl_add: store_every must be >= 1 or -1");
    if (D.n_basis < 0 || (D.n_basis > 0 && !basis)) return fail(ctx, TDSE_ERR_ARG, "tdse_cyl_add: basis missing");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int s = -1;
    for (int i = 0; i < (int)ctx->cyl.size(); ++i)
        if (!ctx->cyl[i].live) {
            s = i;
            break;
        }
    if (s < 0) {
        if ((int)ctx->cyl.size() >= kSlotCapacity) return fail(ctx, TDSE_ERR_NOMEM, "no free simulation slot");
        ctx->cyl.emplace_back();
        s = (int)ctx->cyl.size() - 1;
    }
    CylHost& H = ctx->cyl[s];
    H.blocks.clear();
    CylSim& S = H.dev;
    S = CylSim{};
    S.nz = D.z_points;
    S.nr = D.rho_points;
    S.Pz = Pz;
    S.Mz = Mz;
    S.npz = Pz * Mz;
    S.Cr = Cr;
    S.Mr = Mr;
    S.nrpad = Cr * Mr;
    S.lanes = lanes;
    S.n_groups = S.npz / lanes;
    S.N = D.n_steps;
    S.store_every = D.store_every;
    S.n_basis = D.n_basis;
    S.zp = D.z_prefactor;
    S.rp = D.rho_prefactor;
    S.kc = D.coulomb;
    S.q = D.test_charge;
    S.delta = D.delta;
    S.has_mask = D.has_mask;
    S.m_inner = D.mask_inner;
    S.m_outer = D.mask_outer;
    S.m_inv_s = D.mask_smoothness > 0 ? 1.0 / D.mask_smoothness : 1.0;
    S.m_roots = -1;
    for (int k = 0; k <= 10; ++k)
        if (D.mask_smoothness == (double)(1 << k)) S.m_roots = k;
    S.z0 = D.z_first;
    S.dz = D.z_step;
    S.rho0 = D.rho_first;
    S.drho = D.rho_step;
    S.step = 0;
    S.active = 1;
    S.phase = 1;
    S.hold = 0;
    H.n_data = n_data_points(S.N, S.store_every);
    const size_t pts = (size_t)S.nrpad * S.npz;
    double2 *g, *oip, *bas = nullptr;
    double *dzt, *drt, *droff, *dfs, *dts, *part, *on;
    int rc = 0;
    if ((rc = copy_to_device(ctx, H.blocks, reinterpret_cast<const double2*>(g0), pts, &g)) ||
        (rc = copy_to_device(ctx, H.blocks, zt, (size_t)S.npz, &dzt)) ||
        (rc = copy_to_device(ctx, H.blocks, rt, (size_t)S.nrpad, &drt)) ||
        (rc = copy_to_device(ctx, H.blocks, roff, (size_t)S.nrpad, &droff)) ||
        (rc = copy_to_device(ctx, H.blocks, fs, (size_t)S.N, &dfs)) ||
        (rc = copy_to_device(ctx, H.blocks, ts, (size_t)S.N, &dts)) ||
        (rc = copy_to_device<double>(ctx, H.blocks, nullptr, (size_t)S.n_groups * (1 + 2 * S.n_basis), &part)) ||
        (rc = copy_to_device<double>(ctx, H.blocks, nullptr, (size_t)H.n_data, &on)) ||
        (rc = copy_to_device<double2>(ctx, H.blocks, nullptr, (size_t)H.n_data * std::max(1, S.n_basis), &oip)) ||
        (S.n_basis > 0 &&
         (rc = copy_to_device(ctx, H.blocks, reinterpret_cast<const double2*>(basis), (size_t)S.n_basis * pts, &bas)))) {
        (void)hipStreamSynchronize(ctx->stream);
        free_blocks(ctx, H.blocks);
        return rc;
    }
    S.g = g;
    S.zt = dzt;
    S.rt = drt;
    S.roff = droff;
    S.fs = dfs;
    S.ts = dts;
    S.part = part;
    S.out_norm = on;
    S.out_ip = oip;
    S.basis = bas;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->d_cyl + s, &S, sizeof(CylSim), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // host buffers may be released after return
    H.live = true;
    ctx->dirty = true;
    ctx->cyl_pending = true;
    *slot = s;
    return TDSE_OK;
}

int tdse_set_hold(tdse_ctx* ctx, int32_t kind, int32_t slot, int32_t hold) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (kind == 0 || kind == 1) return TDSE_OK;  // SH and Line states can be made exact at any step
    if (kind != 2) return fail(ctx, TDSE_ERR_ARG, "unknown simulation kind");
    if (slot < 0 || slot >= (int)ctx->cyl.size() || !ctx->cyl[slot].live) return fail(ctx, TDSE_ERR_ARG, "bad Cyl slot");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const int h = hold ? 1 : 0;
    ctx->cyl[slot].dev.hold = h;
    HIP_TRY(ctx, hipMemcpy(reinterpret_cast<char*>(ctx->d_cyl + slot) + offsetof(CylSim, hold), &h, sizeof(int),
                           hipMemcpyHostToDevice));
    return TDSE_OK;
}

int tdse_run(tdse_ctx* ctx, int32_t n_steps, int32_t use_graph) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (n_steps < 0) return fail(ctx, TDSE_ERR_ARG, "tdse_run: n_steps must be >= 0");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc;
    if (ctx->dirty && (rc = rebuild_tables(ctx))) return rc;
    const bool graph = use_graph && !ctx->profiling;
    bool hold_any = false;
    for (const auto& h : ctx->cyl)
        if (h.live && h.dev.hold) hold_any = true;
    ctx->ev_used = 0;
    ctx->ev_pairs.clear();
    // SH ADI simulations may run ahead into the next step on every replay of this run but
    // the last, so that the run ends with every simulation at an ordinary step boundary
    // (mask pending) which the host can materialise, download or checkpoint.
    const bool ahead = !ctx->sh_adi_buckets.empty() && n_steps > 1;
    HIP_TRY(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
    if (!ctx->sh_adi_buckets.empty())
        HIP_TRY(ctx, hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(ctx->d_ctl), ahead ? 1 : 0, 1, ctx->stream));
    for (int i = 0; i < n_steps;) {
        // the pre-sweep is needed for simulations held at exact states, and for
        // one step after a cylindrical simulation was added or uploaded
        int done = 1;
        const int left = n_steps - i - (ahead ? 1 : 0);  // steps before the last one of a run with run-ahead
        if (ahead && left == 0)
            HIP_TRY(ctx, hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(ctx->d_ctl), 0, 1, ctx->stream));
        if (ctx->cyl_pending && !hold_any) {
            if ((rc = launch_step(ctx, true))) return rc;
        } else if (graph) {
            if ((rc = ensure_graph(ctx, hold_any))) return rc;
            if (kGraphSteps > 1 && left >= kGraphSteps) {
                HIP_TRY(ctx, hipGraphLaunch(ctx->exec_k, ctx->stream));
                done = kGraphSteps;
            } else {
                HIP_TRY(ctx, hipGraphLaunch(ctx->exec, ctx->stream));
            }
        } else if ((rc = launch_step(ctx, hold_any))) {
            return rc;
        }
        ctx->cyl_pending = false;
        i += done;
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev_stop));
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    ctx->last_run_ms = ms;
    if (ctx->profiling) {
        for (auto& pr : ctx->ev_pairs) {
            float t = 0.f;
            HIP_TRY(ctx, hipEventElapsedTime(&t, ctx->ev_pool[pr.second], ctx->ev_pool[pr.second + 1]));
            ctx->prof_ms[pr.first] += t;
            ctx->prof_n[pr.first] += 1;
        }
    }
    return TDSE_OK;
}

int tdse_line_run(tdse_ctx* ctx, int32_t n_steps) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (n_steps < 0) return fail(ctx, TDSE_ERR_ARG, "tdse_line_run: n_steps must be >= 0");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc;
    if (ctx->dirty && (rc = rebuild_tables(ctx))) return rc;
    HIP_TRY(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
    for (auto& b : ctx->line_buckets) dispatch_line(ctx, b, n_steps);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev_stop));
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    ctx->last_run_ms = ms;
    return TDSE_OK;
}

int tdse_sim_state(tdse_ctx* ctx, int32_t kind, int32_t slot, int32_t* step, int32_t* n_steps, int32_t* n_data) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    int st = 0;
    if (kind == 0) {
        if (slot < 0 || slot >= (int)ctx->sh.size() || !ctx->sh[slot].live) return fail(ctx, TDSE_ERR_ARG, "bad SH slot");
        HIP_TRY(ctx, hipMemcpy(&st, reinterpret_cast<char*>(ctx->d_sh + slot) + offsetof(ShSim, step), sizeof(int),
                               hipMemcpyDeviceToHost));
        if (n_steps) *n_steps = ctx->sh[slot].dev.N;
        if (n_data) *n_data = ctx->sh[slot].n_data;
    } else if (kind == 1) {
        if (slot < 0 || slot >= (int)ctx->line.size() || !ctx->line[slot].live)
            return fail(ctx, TDSE_ERR_ARG, "bad Line slot");
        HIP_TRY(ctx, hipMemcpy(&st, reinterpret_cast<char*>(ctx->d_line + slot) + offsetof(LineSim, step), sizeof(int),
                               hipMemcpyDeviceToHost));
        if (n_steps) *n_steps = ctx->line[slot].dev.N;
        if (n_data) *n_data = ctx->line[slot].n_data;
    } else if (kind == 2) {
        if (slot < 0 || slot >= (int)ctx->cyl.size() || !ctx->cyl[slot].live) return fail(ctx, TDSE_ERR_ARG, "bad Cyl slot");
        HIP_TRY(ctx, hipMemcpy(&st, reinterpret_cast<char*>(ctx->d_cyl + slot) + offsetof(CylSim, step), sizeof(int),
                               hipMemcpyDeviceToHost));
        if (n_steps) *n_steps = ctx->cyl[slot].dev.N;
        if (n_data) *n_data = ctx->cyl[slot].n_data;
    } else {
        return fail(ctx, TDSE_ERR_ARG, "unknown simulation kind");
    }
    if (step) *step = st;
    return TDSE_OK;
}

int tdse_download(tdse_ctx* ctx, int32_t kind, int32_t slot, double* g, double* norms, double* lnorms, double* ips) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (kind == 0) {
        if (slot < 0 || slot >= (int)ctx->sh.size() || !ctx->sh[slot].live) return fail(ctx, TDSE_ERR_ARG, "bad SH slot");
        const ShHost& H = ctx->sh[slot];
        const ShSim& S = H.dev;
        if (g) {
            int st = 0;
            HIP_TRY(ctx, hipMemcpy(&st, reinterpret_cast<char*>(ctx->d_sh + slot) + offsetof(ShSim, step), sizeof(int),
                                   hipMemcpyDeviceToHost));
            HIP_TRY(ctx, hipMemcpy(g, sh_state(S, st), (size_t)S.L * S.npad * sizeof(double2), hipMemcpyDeviceToHost));
        }
        if (norms) HIP_TRY(ctx, hipMemcpy(norms, S.out_norm, H.n_data * sizeof(double), hipMemcpyDeviceToHost));
        if (lnorms) HIP_TRY(ctx, hipMemcpy(lnorms, S.out_lnorm, H.n_data * sizeof(double), hipMemcpyDeviceToHost));
        if (ips && S.n_basis > 0)
            HIP_TRY(ctx, hipMemcpy(ips, S.out_ip, (size_t)H.n_data * S.n_basis * sizeof(double2), hipMemcpyDeviceToHost));
    } else if (kind == 1) {
        if (slot < 0 || slot >= (int)ctx->line.size() || !ctx->line[slot].live)
            return fail(ctx, TDSE_ERR_ARG, "bad Line slot");
        const LineHost& H = ctx->line[slot];
        const LineSim& S = H.dev;
        if (g) HIP_TRY(ctx, hipMemcpy(g, S.g, (size_t)S.npad * sizeof(double2), hipMemcpyDeviceToHost));
        if (norms) HIP_TRY(ctx, hipMemcpy(norms, S.out_norm, H.n_data * sizeof(double), hipMemcpyDeviceToHost));
        if (ips && S.n_basis > 0)
            HIP_TRY(ctx, hipMemcpy(ips, S.out_ip, (size_t)H.n_data * S.n_basis * sizeof(double2), hipMemcpyDeviceToHost));
    } else if (kind == 2) {
        if (slot < 0 || slot >= (int)ctx->cyl.size() || !ctx->cyl[slot].live) return fail(ctx, TDSE_ERR_ARG, "bad Cyl slot");
        const CylHost& H = ctx->cyl[slot];
        const CylSim& S = H.dev;
        if (g) HIP_TRY(ctx, hipMemcpy(g, S.g, (size_t)S.nrpad * S.npz * sizeof(double2), hipMemcpyDeviceToHost));
        if (norms) HIP_TRY(ctx, hipMemcpy(norms, S.out_norm, H.n_data * sizeof(double), hipMemcpyDeviceToHost));
        if (ips && S.n_basis > 0)
            HIP_TRY(ctx, hipMemcpy(ips, S.out_ip, (size_t)H.n_data * S.n_basis * sizeof(double2), hipMemcpyDeviceToHost));
    } else {
        return fail(ctx, TDSE_ERR_ARG, "unknown simulation kind");
    }
    return TDSE_OK;
}

int tdse_download_norm_by_l(tdse_ctx* ctx, int32_t slot, double* out) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (!out) return fail(ctx, TDSE_ERR_ARG, "out is NULL");
    if (slot < 0 || slot >= (int)ctx->sh.size() || !ctx->sh[slot].live) return fail(ctx, TDSE_ERR_ARG, "bad SH slot");
    const ShHost& H = ctx->sh[slot];
    if (!H.dev.norm_by_l) return fail(ctx, TDSE_ERR_STATE, "simulation was added without norm_by_l");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipMemcpy(out, H.dev.out_nl, (size_t)H.n_data * H.dev.L * sizeof(double), hipMemcpyDeviceToHost));
    return TDSE_OK;
}

int tdse_upload_state(tdse_ctx* ctx, int32_t kind, int32_t slot, const double* g, int32_t step) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (!g) return fail(ctx, TDSE_ERR_ARG, "g is NULL");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (kind == 0) {
        if (slot < 0 || slot >= (int)ctx->sh.size() || !ctx->sh[slot].live) return fail(ctx, TDSE_ERR_ARG, "bad SH slot");
        const ShSim& S = ctx->sh[slot].dev;
        if (step < 0 || step > S.N + 1) return fail(ctx, TDSE_ERR_ARG, "step out of range");
        HIP_TRY(ctx, hipMemcpy(sh_state(S, step), g, (size_t)S.L * S.npad * sizeof(double2), hipMemcpyHostToDevice));
        const int st[2] = {step, 1};  // the uploaded g is the reference g after `step` steps
        HIP_TRY(ctx, hipMemcpy(reinterpret_cast<char*>(ctx->d_sh + slot) + offsetof(ShSim, step), &st[0], sizeof(int),
                               hipMemcpyHostToDevice));
        HIP_TRY(ctx, hipMemcpy(reinterpret_cast<char*>(ctx->d_sh + slot) + offsetof(ShSim, phase), &st[1], sizeof(int),
                               hipMemcpyHostToDevice));
        HIP_TRY(ctx, hipMemcpy(reinterpret_cast<char*>(ctx->d_sh + slot) + offsetof(ShSim, flags), &kNoFlags,
                               sizeof(int), hipMemcpyHostToDevice));
    } else if (kind == 1) {
        if (slot < 0 || slot >= (int)ctx->line.size() || !ctx->line[slot].live)
            return fail(ctx, TDSE_ERR_ARG, "bad Line slot");
        const LineSim& S = ctx->line[slot].dev;
        if (step < 0 || step > S.N + 1) return fail(ctx, TDSE_ERR_ARG, "step out of range");
        HIP_TRY(ctx, hipMemcpy(S.g, g, (size_t)S.npad * sizeof(double2), hipMemcpyHostToDevice));
        HIP_TRY(ctx, hipMemcpy(reinterpret_cast<char*>(ctx->d_line + slot) + offsetof(LineSim, step), &step, sizeof(int),
                               hipMemcpyHostToDevice));
        HIP_TRY(ctx, hipMemcpy(reinterpret_cast<char*>(ctx->d_line + slot) + offsetof(LineSim, flags), &kNoFlags,
                               sizeof(int), hipMemcpyHostToDevice));
    } else if (kind == 2) {
        if (slot < 0 || slot >= (int)ctx->cyl.size() || !ctx->cyl[slot].live) return fail(ctx, TDSE_ERR_ARG, "bad Cyl slot");
        const CylSim& S = ctx->cyl[slot].dev;
        if (step < 0 || step > S.N) return fail(ctx, TDSE_ERR_ARG, "step out of range");
        HIP_TRY(ctx, hipMemcpy(S.g, g, (size_t)S.nrpad * S.npz * sizeof(double2), hipMemcpyHostToDevice));
        const int st[2] = {step, 1};  // the uploaded g is the reference g after `step` steps
        char* base = reinterpret_cast<char*>(ctx->d_cyl + slot);
        HIP_TRY(ctx, hipMemcpy(base + offsetof(CylSim, step), &st[0], sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(ctx, hipMemcpy(base + offsetof(CylSim, phase), &st[1], sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(ctx, hipMemcpy(base + offsetof(CylSim, flags), &kNoFlags, sizeof(int), hipMemcpyHostToDevice));
        ctx->cyl_pending = true;
    } else {
        return fail(ctx, TDSE_ERR_ARG, "unknown simulation kind");
    }
    return TDSE_OK;
}

int tdse_remove(tdse_ctx* ctx, int32_t kind, int32_t slot) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const int zero = 0;
    if (kind == 0) {
        if (slot < 0 || slot >= (int)ctx->sh.size() || !ctx->sh[slot].live) return fail(ctx, TDSE_ERR_ARG, "bad SH slot");
        HIP_TRY(ctx, hipMemcpy(reinterpret_cast<char*>(ctx->d_sh + slot) + offsetof(ShSim, active), &zero, sizeof(int),
                               hipMemcpyHostToDevice));
        free_blocks(ctx, ctx->sh[slot].blocks);
        ctx->sh[slot].live = false;
    } else if (kind == 1) {
        if (slot < 0 || slot >= (int)ctx->line.size() || !ctx->line[slot].live)
            return fail(ctx, TDSE_ERR_ARG, "bad Line slot");
        HIP_TRY(ctx, hipMemcpy(reinterpret_cast<char*>(ctx->d_line + slot) + offsetof(LineSim, active), &zero,
                               sizeof(int), hipMemcpyHostToDevice));
        free_blocks(ctx, ctx->line[slot].blocks);
        ctx->line[slot].live = false;
    } else if (kind == 2) {
        if (slot < 0 || slot >= (int)ctx->cyl.size() || !ctx->cyl[slot].live) return fail(ctx, TDSE_ERR_ARG, "bad Cyl slot");
        HIP_TRY(ctx, hipMemcpy(reinterpret_cast<char*>(ctx->d_cyl + slot) + offsetof(CylSim, active), &zero, sizeof(int),
                               hipMemcpyHostToDevice));
        free_blocks(ctx, ctx->cyl[slot].blocks);
        ctx->cyl[slot].live = false;
    } else {
        return fail(ctx, TDSE_ERR_ARG, "unknown simulation kind");
    }
    ctx->dirty = true;
    return TDSE_OK;
}

int tdse_sim_flags(tdse_ctx* ctx, int32_t kind, int32_t first_slot, int32_t n_slots, int32_t* flags) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (!flags || first_slot < 0 || n_slots < 0) return fail(ctx, TDSE_ERR_ARG, "tdse_sim_flags: bad arguments");
    size_t count, pitch, off;
    const char* base;
    if (kind == 0) {
        count = ctx->sh.size(), pitch = sizeof(ShSim), off = offsetof(ShSim, flags);
        base = reinterpret_cast<const char*>(ctx->d_sh);
    } else if (kind == 1) {
        count = ctx->line.size(), pitch = sizeof(LineSim), off = offsetof(LineSim, flags);
        base = reinterpret_cast<const char*>(ctx->d_line);
    } else if (kind == 2) {
        count = ctx->cyl.size(), pitch = sizeof(CylSim), off = offsetof(CylSim, flags);
        base = reinterpret_cast<const char*>(ctx->d_cyl);
    } else {
        return fail(ctx, TDSE_ERR_ARG, "unknown simulation kind");
    }
    if ((size_t)first_slot + (size_t)n_slots > count) return fail(ctx, TDSE_ERR_ARG, "tdse_sim_flags: slot range");
    if (n_slots == 0) return TDSE_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpy2DAsync(flags, sizeof(int32_t), base + (size_t)first_slot * pitch + off, pitch,
                                  sizeof(int32_t), (size_t)n_slots, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return TDSE_OK;
}

int tdse_materialize(tdse_ctx* ctx, int32_t kind, int32_t slot) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (kind == 1) {
        if (slot < 0 || slot >= (int)ctx->line.size() || !ctx->line[slot].live)
            return fail(ctx, TDSE_ERR_ARG, "bad Line slot");
        return TDSE_OK;  // a LineMesh always holds the reference g between launches
    }
    if (kind == 2) {
        if (slot < 0 || slot >= (int)ctx->cyl.size() || !ctx->cyl[slot].live) return fail(ctx, TDSE_ERR_ARG, "bad Cyl slot");
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        int st[2] = {0, 0};  // step, phase
        char* base = reinterpret_cast<char*>(ctx->d_cyl + slot);
        HIP_TRY(ctx, hipMemcpy(&st[0], base + offsetof(CylSim, step), sizeof(int), hipMemcpyDeviceToHost));
        HIP_TRY(ctx, hipMemcpy(&st[1], base + offsetof(CylSim, phase), sizeof(int), hipMemcpyDeviceToHost));
        if (st[1] == 1) return TDSE_OK;
        return fail(ctx, TDSE_ERR_STATE,
                    "cylindrical state is between steps (explicit rho half applied); call tdse_set_hold before running "
                    "to stop at exact states");
    }
    if (kind != 0) return fail(ctx, TDSE_ERR_ARG, "unknown simulation kind");
    if (slot < 0 || slot >= (int)ctx->sh.size() || !ctx->sh[slot].live) return fail(ctx, TDSE_ERR_ARG, "bad SH slot");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc;
    if (ctx->dirty && (rc = rebuild_tables(ctx))) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    int st[2] = {0, 0};  // step, phase
    char* base = reinterpret_cast<char*>(ctx->d_sh + slot);
    HIP_TRY(ctx, hipMemcpy(&st[0], base + offsetof(ShSim, step), sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(&st[1], base + offsetof(ShSim, phase), sizeof(int), hipMemcpyDeviceToHost));
    const ShHost& H = ctx->sh[slot];
    if (st[1] == 1 || st[0] == 0 || st[0] > H.dev.N) return TDSE_OK;
    const int nw = H.wave_count;
    if (H.vel)
        hipLaunchKernelGGL((k_sh_lphase<false, true>), dim3((nw * TDSE_WAVE + 255) / 256), dim3(256), 0, ctx->stream,
                           ctx->d_sh, ctx->waves.ptr + H.wave_first, nw, ctx->d_clv, 1);
    else
        hipLaunchKernelGGL((k_sh_lphase<false, false>), dim3((nw * TDSE_WAVE + 255) / 256), dim3(256), 0,
                           ctx->stream, ctx->d_sh, ctx->waves.ptr + H.wave_first, nw, ctx->d_cl, 1);
    HIP_TRY(ctx, hipGetLastError());
    // the end-only L-phase wrote sh_state(n + 1): swap the buffers so the exact
    // state is sh_state(n) for the unchanged step counter
    ShSim& S = ctx->sh[slot].dev;
    std::swap(S.g, S.g2);
    const int one = 1;
    HIP_TRY(ctx, hipMemcpyAsync(base + offsetof(ShSim, g), &S.g, sizeof(S.g), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(base + offsetof(ShSim, g2), &S.g2, sizeof(S.g2), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(base + offsetof(ShSim, phase), &one, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return TDSE_OK;
}

int tdse_last_run_ms(tdse_ctx* ctx, double* ms) {
    if (!ctx || !ms) return fail(ctx, TDSE_ERR_ARG, "NULL argument");
    *ms = ctx->last_run_ms;
    return TDSE_OK;
}

int tdse_profile(tdse_ctx* ctx, int32_t enable) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    ctx->profiling = enable != 0;
    for (int i = 0; i < 6; ++i) {
        ctx->prof_ms[i] = 0.0;
        ctx->prof_n[i] = 0;
    }
    return TDSE_OK;
}

int tdse_profile_read(tdse_ctx* ctx, double* lphase_ms, int64_t* lphase_launches, double* rphase_ms,
                      int64_t* rphase_launches) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (lphase_ms) *lphase_ms = ctx->prof_ms[0];
    if (lphase_launches) *lphase_launches = ctx->prof_n[0];
    if (rphase_ms) *rphase_ms = ctx->prof_ms[1];
    if (rphase_launches) *rphase_launches = ctx->prof_n[1];
    return TDSE_OK;
}

int tdse_profile_kernel(tdse_ctx* ctx, int32_t kernel, double* ms, int64_t* launches) {
    if (!ctx) return fail(nullptr, TDSE_ERR_ARG, "ctx is NULL");
    if (kernel < 0 || kernel > 5) return fail(ctx, TDSE_ERR_ARG, "kernel must be 0..5");
    if (ms) *ms = ctx->prof_ms[kernel];
    if (launches) *launches = ctx->prof_n[kernel];
    return TDSE_OK;
}

}  // extern "C"
