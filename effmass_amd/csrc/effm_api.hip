// effm_api.hip — host side of the C-ABI declared in include/effm_hpsi.h.
//
// Owns the device-resident cell (V_loc, beta tables, channels, FFT plans)
// and k-point data (G-sphere, plane index, projector matrices), and
// sequences the kernels of k_local.hip / k_nonlocal.hip on one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <iterator>
#include <mutex>
#include <new>

#include "effm_hpsi.h"
#include "effm_internal.hpp"

struct effm_ctx {
    effm::Ctx c;
};

namespace effm {

// ---------------------------------------------------------------------------
//  Static scratch of effm_overlap_noalloc: the split-K partial sums of an
//  overlap, at most kOverlapPartialCap bytes, in a device array that every
//  GPU the library runs on holds until the process ends.  In host code
//  `scratch` names a host-side stand-in, not the device memory: the device
//  address comes from hipGetSymbolAddress, once per device with that device
//  current (at effm_ctx_create), and only that address is passed to kernels.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kOverlapPartialCap];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};
static hipError_t g_scratch_rc[kMaxDevices] = {};
static bool g_scratch_done[kMaxDevices] = {};

// The calling thread has `device` current.
static void scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_done[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_rc[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
    g_scratch_done[device] = true;
}

// Device address of this device's scratch, or NULL (with the reason in c.err).
static void* scratch_ptr(Ctx& c) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (c.device < 0 || c.device >= kMaxDevices || !g_scratch_done[c.device]) {
        c.err = "static scratch not set up for this device";
        return nullptr;
    }
    if (!g_scratch_ptr[c.device])
        c.err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(g_scratch_rc[c.device]);
    return g_scratch_ptr[c.device];
}

// ---------------------------------------------------------------------------
//  small helpers
// ---------------------------------------------------------------------------
static int fail(Ctx& c, int code, const std::string& msg) {
    c.err = msg;
    return code;
}

static int hip_fail(Ctx& c, hipError_t e, const char* where) {
    c.err = std::string(where) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory) return kErrNoMem;
    return kErrHip;
}

#define HIP_TRY(expr, where)                          \
    do {                                              \
        hipError_t _e = (expr);                       \
        if (_e != hipSuccess) return hip_fail(c, _e, where); \
    } while (0)

// Device buffers are stream-ordered (hipMallocAsync / hipFreeAsync on the
// context's stream): a free waits only for this stream's earlier work, not
// for the whole device, so contexts that come and go (one per served query)
// do not stall the others.
static void dfree(Ctx& c, DevBuf& b) {
    if (b.ptr) {
        (void)hipFreeAsync(b.ptr, c.stream);
        c.used -= b.bytes;
        ++c.epoch;
    }
    b.ptr = nullptr;
    b.bytes = 0;
}

static int dalloc(Ctx& c, DevBuf& b, size_t bytes, const char* what) {
    dfree(c, b);
    if (bytes == 0) return kOk;
    if (c.budget && c.used + bytes > c.budget)
        return fail(c, kErrNoMem, std::string("HBM budget exceeded allocating ") + what + " (" +
                                      std::to_string(bytes) + " B, in use " + std::to_string(c.used) +
                                      " B, budget " + std::to_string(c.budget) + " B)");
    hipError_t e = hipMallocAsync(&b.ptr, bytes, c.stream);
    // the buffer is used right away, also by synchronous copies
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        return fail(c, kErrNoMem, std::string("hipMallocAsync failed for ") + what + ": " + hipGetErrorString(e));
    }
    b.bytes = bytes;
    c.used += bytes;
    ++c.epoch;
    return kOk;
}

static int ensure(Ctx& c, DevBuf& b, size_t bytes, const char* what) {
    if (b.bytes >= bytes && b.ptr) return kOk;
    return dalloc(c, b, bytes, what);
}

template <typename V>
static int upload(Ctx& c, DevBuf& b, const std::vector<V>& h, const char* what) {
    int rc = dalloc(c, b, h.size() * sizeof(V), what);
    if (rc) return rc;
    if (!h.empty()) HIP_TRY(hipMemcpy(b.ptr, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice), what);
    return kOk;
}

// Radix sequence for an axis: fewest stages (each stage is one pass over
// LDS), radices from the in-register DFTs of fft_lds.hpp; ties go to the
// sequence whose largest radix is smallest (fewer registers).
static bool factor_axis(int n, FftAxis& ax) {
    static const int radices[] = {13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2};
    ax.n = n;
    ax.ns = 0;
    if (n == 1) return true;
    std::vector<int> best(n + 1, 1 << 20), pick(n + 1, 0), worst(n + 1, 0);
    best[1] = 0;
    for (int m = 2; m <= n; ++m) {
        if (n % m) continue;
        for (int p : radices) {
            if (m % p || best[m / p] >= (1 << 20)) continue;
            int st = best[m / p] + 1;
            int wr = std::max(worst[m / p], p);
            if (st < best[m] || (st == best[m] && wr < worst[m])) {
                best[m] = st;
                pick[m] = p;
                worst[m] = wr;
            }
        }
    }
    if (best[n] >= (1 << 20) || best[n] > kMaxStages) return false;
    for (int m = n; m > 1; m /= pick[m]) ax.p[ax.ns++] = pick[m];
    return true;
}

// perm(j): position of input j for the in-place DIT (= position of output
// j for the DIF), the mixed-radix digit reversal of fft_lds.hpp.
static std::vector<int> fft_perm(const FftAxis& ax) {
    std::vector<int> L(ax.ns + 1, 1);
    for (int t = 0; t < ax.ns; ++t) L[t + 1] = L[t] * ax.p[t];
    std::vector<int> P(ax.n);
    for (int j = 0; j < ax.n; ++j) {
        int jj = j, pos = 0;
        for (int t = ax.ns - 1; t >= 0; --t) {
            int d = jj % ax.p[t];
            jj /= ax.p[t];
            pos += d * L[t];
        }
        P[j] = pos;
    }
    return P;
}

static void twiddles(int n, std::vector<double2>& w64, std::vector<float2>& w32) {
    w64.resize(n);
    w32.resize(n);
    for (int k = 0; k < n; ++k) {
        long double a = 2.0L * (long double)k / (long double)n;   // angle / pi
        long double pi = 3.141592653589793238462643383279502884L;
        double cr = (double)cosl(pi * a), ci = (double)(-sinl(pi * a));
        w64[k] = double2{cr, ci};
        w32[k] = float2{(float)cr, (float)ci};
    }
}

void timer_begin(Ctx& c) {
    if (!c.timing || c.capturing) return;
    hipEvent_t e;
    if (!c.event_pool.empty()) {
        e = c.event_pool.back();
        c.event_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        c.cur_start = nullptr;
        return;
    }
    (void)hipEventRecord(e, c.stream);
    c.cur_start = e;
}

void timer_end(Ctx& c, const char* name) {
    if (!c.timing || c.capturing || !c.cur_start) return;
    hipEvent_t e;
    if (!c.event_pool.empty()) {
        e = c.event_pool.back();
        c.event_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        return;
    }
    (void)hipEventRecord(e, c.stream);
    c.pending.push_back({name, c.cur_start, e});
    c.cur_start = nullptr;
}

static void timer_resolve(Ctx& c) {
    for (auto& p : c.pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto& t = c.timers[p.name];
            t.ms += ms;
            t.count += 1;
        }
        c.event_pool.push_back(p.a);
        c.event_pool.push_back(p.b);
    }
    c.pending.clear();
}

static void release_kpt(Ctx& c, KptDev& k) {
    for (DevBuf* b : {&k.K, &k.kin, &k.gidx, &k.plane_x, &k.x_to_plane, &k.plane_ptr, &k.ent_ig,
                      &k.ent_pos, &k.row_ptr, &k.rows, &k.ent_rowptr, &k.P64, &k.P32, &k.dP64, &k.dP32})
        dfree(c, *b);
    k = KptDev{};
}

static void release_cell(Ctx& c) {
    for (auto& k : c.kpts) release_kpt(c, k);
    c.kpts.clear();
    for (DevBuf* b : {&c.vloc64, &c.vloc32, &c.tw1_64, &c.tw2_64, &c.tw3_64, &c.tw1_32, &c.tw2_32,
                      &c.tw3_32, &c.perm1, &c.tau, &c.btab, &c.dbtab, &c.tab_info, &c.chan_dev,
                      &c.lam64})
        dfree(c, *b);
    c.chans.clear();
    c.tinfo.clear();
    c.have_cell = false;
}

static int check_kpt(Ctx& c, int ikpt) {
    if (!c.have_cell) return fail(c, kErrRuntime, "H|psi> requires a cell (effm_cell_upload)");
    if (ikpt < 0 || ikpt >= (int)c.kpts.size() || !c.kpts[ikpt].valid)
        return fail(c, kErrOutOfRange, "Invalid k-point index: " + std::to_string(ikpt));
    return kOk;
}

// Bytes of FFT plane workspace per band.
template <typename T>
static size_t work_per_band(const Ctx& c, const KptDev& k) {
    return (size_t)k.nplanes * ws_plane<T>(c.n2, c.n3) * sizeof(cx_t<T>);
}

static constexpr size_t kWorkCap = size_t(8) << 30;   // FFT workspace cap (bytes)

template <typename T>
static int ensure_proj(Ctx& c, KptDev& k, bool grad) {
    using C = cx_t<T>;
    if (c.nch() == 0) return kOk;
    const bool f64 = sizeof(T) == 8;
    DevBuf& P = f64 ? k.P64 : k.P32;
    DevBuf& dP = f64 ? k.dP64 : k.dP32;
    const size_t pbytes = (size_t)c.nch() * k.npw * sizeof(C);
    bool need = !P.ptr || (grad && !dP.ptr);
    if (!need) return kOk;
    if (grad && !c.have_dbeta)
        return fail(c, kErrRuntime, "dH/dk requires dbeta/dq tables for every species");
    int rc;
    if (!P.ptr && (rc = dalloc(c, P, pbytes, "projectors"))) return rc;
    if (grad && !dP.ptr && (rc = dalloc(c, dP, 3 * pbytes, "projector k-gradients"))) return rc;
    HIP_TRY(launch_build_proj<T>(c, k, grad), "build projectors");
    return kOk;
}

// Workspace cap: kWorkCap (EFFM_WORK_MB overrides), and within the HBM
// budget next to everything else the context holds.
static size_t work_cap(const Ctx& c) {
    size_t cap = kWorkCap;
    if (const char* e = std::getenv("EFFM_WORK_MB")) cap = std::max<size_t>(1, std::strtoull(e, nullptr, 10)) << 20;
    if (c.budget) {
        const size_t others = c.used - c.work.bytes;
        cap = std::min(cap, c.budget > others ? (size_t)(c.budget - others) : (size_t)0);
    }
    return cap;
}

// H|psi> for nseg segments (k-points), kMaxSeg per launch.  launch = false
// only allocates (projector matrices, workspace, partial sums), so that a
// following launch = true pass enqueues kernels and nothing else.
template <typename T>
static int run_hpsi(Ctx& c, const ApplySeg* segs, int nseg, int nb, bool acc, bool launch) {
    using C = cx_t<T>;
    int rc;
    for (int i = 0; i < nseg; ++i)
        if ((rc = ensure_proj<T>(c, *const_cast<KptDev*>(segs[i].k), false))) return rc;
    const size_t cap = work_cap(c);
    for (int g0 = 0; g0 < nseg; g0 += kMaxSeg) {
        const int ng = std::min(kMaxSeg, nseg - g0);
        const ApplySeg* sg = segs + g0;
        const size_t wpb = std::max<size_t>(1, local_work_bytes<T>(c, sg, ng, 1));
        const int chunk = (int)std::max<size_t>(1, std::min<size_t>({(size_t)nb, cap / wpb, (size_t)65535}));
        if ((rc = ensure(c, c.work, wpb * chunk, "FFT workspace"))) return rc;
        if (c.nch() > 0) {
            int mn = sg[0].k->npw;
            for (int i = 1; i < ng; ++i) mn = std::min(mn, sg[i].k->npw);
            if ((rc = ensure(c, c.partial, vnl_partial_bytes<T>(c, c.nch(), nb, mn, ng),
                             "projection partial sums")))
                return rc;
            if ((rc = ensure(c, c.cbuf, (size_t)ng * c.nch() * nb * sizeof(C), "projections"))) return rc;
        }
        if (!launch) continue;
        for (int b0 = 0; b0 < nb; b0 += chunk) {
            const int nbc = std::min(chunk, nb - b0);
            HIP_TRY(launch_local<T>(c, sg, ng, b0, nbc, acc, c.work.ptr), "local potential kernels");
        }
        if (c.nch() > 0)
            HIP_TRY(launch_vnl<T>(c, sg, ng, nb, c.partial.ptr, c.partial.bytes, c.cbuf.ptr), "nonlocal kernels");
    }
    return kOk;
}

template <typename T>
// with_h: H|psi> into each segment's hout as well, its nonlocal part from
// the <p|psi> block of the gradient's projection (one projection GEMM for
// both instead of two).
static int run_dhdk(Ctx& c, const ApplySeg* segs, int nseg, int nb, bool acc, bool launch, bool with_h = false) {
    using C = cx_t<T>;
    int rc;
    for (int i = 0; i < nseg; ++i)
        if ((rc = ensure_proj<T>(c, *const_cast<KptDev*>(segs[i].k), true))) return rc;
    const size_t cap = work_cap(c);
    for (int g0 = 0; g0 < nseg; g0 += kMaxSeg) {
        const int ng = std::min(kMaxSeg, nseg - g0);
        const ApplySeg* sg = segs + g0;
        // the local part of H|psi>: segments writing to hout
        ApplySeg hs[kMaxSeg];
        int chunk = 0;
        if (with_h) {
            for (int i = 0; i < ng; ++i) {
                hs[i] = sg[i];
                hs[i].out[0] = sg[i].hout;
                hs[i].out[1] = hs[i].out[2] = nullptr;
            }
            const size_t wpb = std::max<size_t>(1, local_work_bytes<T>(c, hs, ng, 1));
            chunk = (int)std::max<size_t>(1, std::min<size_t>({(size_t)nb, cap / wpb, (size_t)65535}));
            if ((rc = ensure(c, c.work, wpb * chunk, "FFT workspace"))) return rc;
        }
        if (c.nch() > 0) {
            int mn = sg[0].k->npw;
            for (int i = 1; i < ng; ++i) mn = std::min(mn, sg[i].k->npw);
            if ((rc = ensure(c, c.partial, vnl_partial_bytes<T>(c, 4 * c.nch(), nb, mn, ng),
                             "projection partial sums")))
                return rc;
            if ((rc = ensure(c, c.cbuf, (size_t)ng * 4 * c.nch() * nb * sizeof(C), "projections"))) return rc;
        }
        if (!launch) continue;
        if (with_h)
            for (int b0 = 0; b0 < nb; b0 += chunk) {
                const int nbc = std::min(chunk, nb - b0);
                HIP_TRY(launch_local<T>(c, hs, ng, b0, nbc, acc, c.work.ptr), "local potential kernels");
            }
        HIP_TRY(launch_dhdk<T>(c, sg, ng, nb, acc, c.partial.ptr, c.partial.bytes, c.cbuf.ptr, with_h),
                "dH/dk kernels");
    }
    return kOk;
}

// ---------------------------------------------------------------------------
//  hipGraph capture / replay of an apply sequence (EFFM_F_GRAPH)
// ---------------------------------------------------------------------------
static constexpr size_t kMaxGraphs = 16;

static void drop_graphs(Ctx& c) {
    for (auto& g : c.graphs) (void)hipGraphExecDestroy(g.exec);
    c.graphs.clear();
}

static std::vector<int64_t> graph_key(int kind, int prec, bool acc, int nb, const ApplySeg* segs, int nseg) {
    std::vector<int64_t> key = {kind, prec, acc ? 1 : 0, nb, nseg};
    for (int i = 0; i < nseg; ++i) {
        key.push_back((int64_t)(intptr_t)segs[i].k);
        key.push_back((int64_t)(intptr_t)segs[i].psi);
        key.push_back(segs[i].ld);
        for (int a = 0; a < 3; ++a) key.push_back((int64_t)(intptr_t)segs[i].out[a]);
        key.push_back(segs[i].ldo);
    }
    return key;
}

// Runs `enqueue` (a launch = true pass) eagerly, or, with graph = true, as a
// replay of a captured graph of the same launches (captured on first use).
// Allocation happened before (launch = false pass), so the capture records
// kernels only: no allocation, copy or memset nodes.
template <typename F>
static int run_maybe_graph(Ctx& c, bool graph, std::vector<int64_t> key, F&& enqueue) {
    if (!graph) return enqueue();
    if (c.graph_epoch != c.epoch) {
        drop_graphs(c);
        c.graph_epoch = c.epoch;
    }
    key.push_back((int64_t)(intptr_t)c.stream);
    for (auto& g : c.graphs)
        if (g.key == key) {
            g.uses = ++c.graph_clock;
            HIP_TRY(hipGraphLaunch(g.exec, c.stream), "graph launch");
            return kOk;
        }
    if (!c.capture_stream) HIP_TRY(hipStreamCreateWithFlags(&c.capture_stream, hipStreamNonBlocking), "capture stream");
    hipStream_t user = c.stream;
    HIP_TRY(hipStreamBeginCapture(c.capture_stream, hipStreamCaptureModeRelaxed), "begin capture");
    c.stream = c.capture_stream;
    c.capturing = true;
    int rc = enqueue();
    c.capturing = false;
    c.stream = user;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c.capture_stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(c, e, "end capture");
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(c, e, "graph instantiate");
    if (c.graphs.size() >= kMaxGraphs) {
        auto lru = std::min_element(c.graphs.begin(), c.graphs.end(),
                                    [](const Ctx::Graph& a, const Ctx::Graph& b) { return a.uses < b.uses; });
        (void)hipGraphExecDestroy(lru->exec);
        c.graphs.erase(lru);
    }
    c.graphs.push_back({std::move(key), exec, ++c.graph_clock});
    HIP_TRY(hipGraphLaunch(exec, c.stream), "graph launch");
    return kOk;
}

// ---------------------------------------------------------------------------
//  autotuner: time each candidate launch shape on this cell / k-point
// ---------------------------------------------------------------------------
// Each knob is timed on the kernels it changes (per-kernel HIP events inside
// an apply), not on the whole apply: the rest of the apply only adds noise
// (whole-apply timing once picked a pencil shape that was slower per kernel).
template <typename T>
static int autotune(Ctx& c, KptDev& k, int nb, int* pencil_bands, int* gemm_wgs) {
    using C = cx_t<T>;
    const size_t bytes = (size_t)nb * k.npw * sizeof(C);
    int rc;
    if ((rc = ensure(c, c.stage_in, bytes, "tuning psi")) || (rc = ensure(c, c.stage_out0, bytes, "tuning hpsi")))
        return rc;
    // any non-zero, non-denormal pattern: timings do not depend on the values
    HIP_TRY(hipMemsetAsync(c.stage_in.ptr, 0x3c, bytes, c.stream), "tuning fill");
    ApplySeg seg{&k, c.stage_in.ptr, k.npw, {c.stage_out0.ptr, nullptr, nullptr}, k.npw};
    Ctx::Tune& tn = c.tune[sizeof(T) == 8 ? 0 : 1];
    // the caller's timing switch and accumulated timers are kept aside
    timer_resolve(c);
    const bool timing = c.timing;
    auto saved = std::move(c.timers);
    c.timers.clear();
    c.timing = true;
    // best of three applies after one warm-up: the summed milliseconds of
    // the named kernel classes (0 when none of them runs on this cell)
    auto measure = [&](std::initializer_list<const char*> names, float* best) -> int {
        *best = 1e30f;
        int r0 = run_hpsi<T>(c, &seg, 1, nb, false, false);   // allocation pass
        if (r0) return r0;
        for (int rep = 0; rep < 4; ++rep) {
            c.timers.clear();
            const int r = run_hpsi<T>(c, &seg, 1, nb, false, true);
            timer_resolve(c);
            if (r) return r;
            double ms = 0.0;
            for (const char* n : names) {
                auto it = c.timers.find(n);
                if (it != c.timers.end()) ms += it->second.ms;
            }
            if (rep > 0) *best = std::min(*best, (float)ms);
        }
        return kOk;
    };
    // the candidate of `choices` (written to *knob) with the least time;
    // the knob keeps its value when its kernels do not run
    auto pick = [&](int* knob, const int* choices, int nchoice, std::initializer_list<const char*> names) -> int {
        float best = 1e30f, t = 0.f;
        int keep = *knob, r = kOk;
        for (int i = 0; i < nchoice; ++i) {
            *knob = choices[i];
            if ((r = measure(names, &t))) break;
            if (t <= 0.f) break;
            if (t < best) best = t, keep = choices[i];
        }
        *knob = keep;
        return r;
    };
    rc = kOk;
    // x-pencil: band group and z width together (one kernel, 3 x 2 shapes)
    if (rc == kOk) {
        float best = 1e30f, t = 0.f;
        int kb = tn.pencil_bands, kz = tn.pencil_z;
        for (int bz : kPencilBandChoices) {
            for (int z : kPencilZChoices) {
                tn.pencil_bands = bz;
                tn.pencil_z = z;
                if ((rc = measure({"xpencil_vloc"}, &t))) break;
                if (t > 0.f && t < best) best = t, kb = bz, kz = z;
            }
            if (rc) break;
        }
        tn.pencil_bands = kb;
        tn.pencil_z = kz;
    }
    if (rc == kOk)
        rc = pick(&tn.gemm_wgs, kGemmWgChoices, (int)std::size(kGemmWgChoices), {"proj_gemm", "proj_reduce"});
    if (rc == kOk && nb > 32)
        rc = pick(&tn.apply_tile, kApplyTileChoices, (int)std::size(kApplyTileChoices), {"proj_apply"});
    tn.tuned = rc == kOk;
    c.timers = std::move(saved);
    c.timing = timing;
    ++c.epoch;    // captured graphs hold the old shapes: recapture
    if (pencil_bands) *pencil_bands = tn.pencil_bands;
    if (gemm_wgs) *gemm_wgs = tn.gemm_wgs;
    return rc;
}

template <typename T>
static int apply_hpsi(Ctx& c, const ApplySeg* segs, int nseg, int nb, bool acc, bool graph, int prec) {
    int rc = run_hpsi<T>(c, segs, nseg, nb, acc, false);
    if (rc) return rc;
    return run_maybe_graph(c, graph, graph ? graph_key(0, prec, acc, nb, segs, nseg) : std::vector<int64_t>{},
                           [&] { return run_hpsi<T>(c, segs, nseg, nb, acc, true); });
}

template <typename T>
static int apply_dhdk(Ctx& c, const ApplySeg* segs, int nseg, int nb, bool acc, bool graph, int prec) {
    int rc = run_dhdk<T>(c, segs, nseg, nb, acc, false);
    if (rc) return rc;
    return run_maybe_graph(c, graph, graph ? graph_key(1, prec, acc, nb, segs, nseg) : std::vector<int64_t>{},
                           [&] { return run_dhdk<T>(c, segs, nseg, nb, acc, true); });
}

template <typename T>
static int apply_hd(Ctx& c, const ApplySeg* segs, int nseg, int nb, bool acc, bool graph, int prec) {
    int rc = run_dhdk<T>(c, segs, nseg, nb, acc, false, true);
    if (rc) return rc;
    std::vector<int64_t> key;
    if (graph) {
        key = graph_key(2, prec, acc, nb, segs, nseg);
        for (int i = 0; i < nseg; ++i) key.push_back((int64_t)(intptr_t)segs[i].hout);
    }
    return run_maybe_graph(c, graph, std::move(key), [&] { return run_dhdk<T>(c, segs, nseg, nb, acc, true, true); });
}

}  // namespace effm

using namespace effm;

// ===========================================================================
//  C-ABI
// ===========================================================================
extern "C" {

int effm_ctx_create(effm_ctx** out, int device, uint64_t hbm_budget) {
    if (!out) return kErrInvalid;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return kErrHip;
    if (device < 0 || device >= ndev) return kErrOutOfRange;
    if (hipSetDevice(device) != hipSuccess) return kErrHip;
    auto* h = new (std::nothrow) effm_ctx;
    if (!h) return kErrNoMem;
    h->c.device = device;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
        h->c.ncu = ncu;
    h->c.budget = hbm_budget;
    h->c.stream = nullptr;   // legacy default stream until effm_set_stream
    scratch_init(device);
    prime_occupancy(h->c);
    *out = h;
    return kOk;
}

int effm_ctx_destroy(effm_ctx* h) {
    if (!h) return kOk;
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    (void)hipStreamSynchronize(c.stream);
    release_cell(c);
    for (DevBuf* b : {&c.work, &c.partial, &c.cbuf, &c.stage_in, &c.stage_out0, &c.stage_out1, &c.stage_out2})
        dfree(c, *b);
    timer_resolve(c);
    for (auto e : c.event_pool) (void)hipEventDestroy(e);
    drop_graphs(c);
    if (c.capture_stream) (void)hipStreamDestroy(c.capture_stream);
    (void)hipStreamSynchronize(c.stream);      // the stream-ordered frees above
    delete h;
    return kOk;
}

int effm_set_stream(effm_ctx* h, void* stream) {
    if (!h) return kErrInvalid;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // buffers are freed in the order of the current stream: finish the old
    // stream's work before switching
    if (s != h->c.stream) (void)hipStreamSynchronize(h->c.stream);
    h->c.stream = s;
    return kOk;
}

const char* effm_last_error(const effm_ctx* h) {
    if (!h) return "null context";
    return h->c.err.c_str();
}

int effm_mem_info(const effm_ctx* h, uint64_t* used, uint64_t* budget) {
    if (!h) return kErrInvalid;
    if (used) *used = h->c.used;
    if (budget) *budget = h->c.budget;
    return kOk;
}

int effm_cell_upload(effm_ctx* h, const effm_cell* cell) {
    if (!h || !cell) return kErrInvalid;
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    if (cell->n1 <= 0 || cell->n2 <= 0 || cell->n3 <= 0)
        return fail(c, kErrInvalid, "FFT grid dimensions must be positive");
    if (!cell->vloc) return fail(c, kErrRuntime, "H|psi> requires VR (vloc is NULL)");
    if (cell->natom < 0 || (cell->natom > 0 && (!cell->species || !cell->tau)))
        return fail(c, kErrRuntime, "H|psi> requires ATOM (species / tau missing)");
    FftAxis a1, a2, a3;
    if (!factor_axis(cell->n1, a1) || !factor_axis(cell->n2, a2) || !factor_axis(cell->n3, a3))
        return fail(c, kErrInvalid,
                    "FFT grid sizes must factor into 2, 3, 5, 7, 11 and 13 (at most " +
                        std::to_string(kMaxStages) + " factors per axis)");
    release_cell(c);
    c.n1 = cell->n1; c.n2 = cell->n2; c.n3 = cell->n3;
    std::memcpy(c.A, cell->lattice, sizeof(c.A));
    const double* A = cell->lattice;
    double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                 A[2] * (A[3] * A[7] - A[4] * A[6]);
    c.omega = std::fabs(det);
    if (!(c.omega > 0)) return fail(c, kErrInvalid, "lattice is singular");
    c.ax1 = a1; c.ax2 = a2; c.ax3 = a3;
    c.hperm1 = fft_perm(a1); c.hperm2 = fft_perm(a2); c.hperm3 = fft_perm(a3);
    int pp1 = 0, pp2 = 0;
    if (plane_ct_pfa(c, &pp1, &pp2)) {
        // prime-factor order of the plane kernels: natural j = (P2 u + P1 b)
        // mod n sits at b P1 + u, i.e. u = (j mod P1) P2^-1 mod P1 and
        // b = (j mod P2) P1^-1 mod P2
        int inv1 = 1, inv2 = 1;
        while ((pp2 * inv1) % pp1 != 1 % pp1) ++inv1;
        while ((pp1 * inv2) % pp2 != 1 % pp2) ++inv2;
        for (int j = 0; j < c.n2; ++j) {
            const int u = ((j % pp1) * inv1) % pp1, b = ((j % pp2) * inv2) % pp2;
            c.hperm2[j] = c.hperm3[j] = b * pp1 + u;
        }
    }
    int rc;
    size_t N = (size_t)c.n1 * c.n2 * c.n3;
    std::vector<double> v64(cell->vloc, cell->vloc + N);
    std::vector<float> v32(N);
    for (size_t i = 0; i < N; ++i) v32[i] = (float)v64[i];
    if ((rc = upload(c, c.vloc64, v64, "vloc"))) return rc;
    if ((rc = upload(c, c.vloc32, v32, "vloc fp32"))) return rc;
    std::vector<double2> w64;
    std::vector<float2> w32;
    twiddles(c.n1, w64, w32);
    if ((rc = upload(c, c.tw1_64, w64, "twiddles")) || (rc = upload(c, c.tw1_32, w32, "twiddles"))) return rc;
    twiddles(c.n2, w64, w32);
    if ((rc = upload(c, c.tw2_64, w64, "twiddles")) || (rc = upload(c, c.tw2_32, w32, "twiddles"))) return rc;
    twiddles(c.n3, w64, w32);
    if ((rc = upload(c, c.tw3_64, w64, "twiddles")) || (rc = upload(c, c.tw3_32, w32, "twiddles"))) return rc;
    if ((rc = upload(c, c.perm1, c.hperm1, "perm"))) return rc;

    // atoms, tables, channels (reference order: species -> atom -> (l, ib) -> m)
    c.natom = cell->natom;
    std::vector<double> tau(cell->tau, cell->tau + 3 * (size_t)cell->natom);
    if ((rc = upload(c, c.tau, tau, "tau"))) return rc;
    const double norm = 1.0 / std::sqrt(c.omega);
    int nq_max = 0;
    bool all_dbeta = true;
    for (int s = 0; s < cell->nspecies; ++s) {
        const effm_species& sp = cell->sp[s];
        if (sp.lmax < 0 || sp.nbeta <= 0) continue;
        nq_max = std::max(nq_max, (int)sp.nq);
        if (!sp.dbeta_tab) all_dbeta = false;
    }
    std::vector<int> tab_of(0);
    std::vector<double> btab, dbtab;
    std::vector<TableDev> tdev;
    c.lmax = -1;
    c.tables_qmax = 1e300;
    std::vector<std::vector<int>> sp_table(cell->nspecies);
    for (int s = 0; s < cell->nspecies; ++s) {
        const effm_species& sp = cell->sp[s];
        if (sp.lmax < 0 || sp.nbeta <= 0) continue;
        if (sp.nq < 4 || !(sp.dq > 0) || !sp.beta_tab || !sp.beta_l || !sp.lambda)
            return fail(c, kErrInvalid, "species " + std::to_string(s) + ": incomplete beta tables");
        if (sp.lmax > 6) return fail(c, kErrInvalid, "lmax > 6 is not supported");
        c.lmax = std::max(c.lmax, (int)sp.lmax);
        for (int ib = 0; ib < sp.nbeta; ++ib) {
            if (ib > 0 && sp.beta_l[ib] < sp.beta_l[ib - 1])
                return fail(c, kErrInvalid, "beta channels must be sorted by l (sortByL)");
            if (sp.beta_l[ib] < 0 || sp.beta_l[ib] > sp.lmax)
                return fail(c, kErrInvalid, "beta_l out of [0, lmax]");
            int row = (int)tdev.size();
            sp_table[s].push_back(row);
            btab.resize((size_t)(row + 1) * nq_max, 0.0);
            std::copy(sp.beta_tab + (size_t)ib * sp.nq, sp.beta_tab + (size_t)(ib + 1) * sp.nq,
                      btab.begin() + (size_t)row * nq_max);
            TableDev t{};
            t.nq = sp.nq;
            t.dq = sp.dq;
            t.qmax = sp.qmax;
            t.norm = norm;
            t.dbeta0 = 0.0;
            if (all_dbeta) {
                dbtab.resize((size_t)(row + 1) * nq_max, 0.0);
                std::copy(sp.dbeta_tab + (size_t)ib * sp.nq, sp.dbeta_tab + (size_t)(ib + 1) * sp.nq,
                          dbtab.begin() + (size_t)row * nq_max);
                t.dbeta0 = sp.dbeta_tab[(size_t)ib * sp.nq] * norm;   // interpolate(q = 0)
            }
            // largest q the 4-point stencil can serve: i0 + 3 < nq
            double qcap = std::min(sp.qmax, sp.dq * (double)(sp.nq - 3) - 1e-12 * sp.dq);
            c.tables_qmax = std::min(c.tables_qmax, qcap);
            tdev.push_back(t);
        }
    }
    c.have_dbeta = all_dbeta;
    c.nq_max = nq_max;
    // channels
    std::vector<ChanDev> chd;
    std::vector<double> lam;
    for (int s = 0; s < cell->nspecies; ++s) {
        const effm_species& sp = cell->sp[s];
        if (sp.lmax < 0 || sp.nbeta <= 0) continue;
        for (int a = 0; a < cell->natom; ++a) {
            if (cell->species[a] != s) continue;
            for (int ib = 0; ib < sp.nbeta; ++ib) {
                double l_ = sp.lambda[ib];
                if (std::fabs(l_) < 1e-15) continue;   // callable.cpp:227
                int l = sp.beta_l[ib];
                for (int m = -l; m <= l; ++m) {
                    c.chans.push_back({a, sp_table[s][ib], l, m, l_, s});
                    chd.push_back({a, sp_table[s][ib], l, m});
                    lam.push_back(l_);
                }
            }
        }
    }
    for (int a = 0; a < cell->natom; ++a)
        if (cell->species[a] < 0 || cell->species[a] >= cell->nspecies)
            return fail(c, kErrOutOfRange, "atom species index out of range");
    if (!tdev.empty()) {
        if ((rc = upload(c, c.btab, btab, "beta tables"))) return rc;
        if (all_dbeta && (rc = upload(c, c.dbtab, dbtab, "dbeta tables"))) return rc;
        if ((rc = upload(c, c.tab_info, tdev, "table info"))) return rc;
    }
    if (!chd.empty()) {
        if ((rc = upload(c, c.chan_dev, chd, "channels"))) return rc;
        if ((rc = upload(c, c.lam64, lam, "lambda"))) return rc;
    }
    c.have_cell = true;
    return kOk;
}

int effm_kpt_upload(effm_ctx* h, int ikpt, int npw, const double* K, const double* kinetic,
                    const int32_t* gidx, const double kfrac[3]) {
    if (!h) return kErrInvalid;
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    if (!c.have_cell) return fail(c, kErrRuntime, "effm_kpt_upload requires a cell");
    if (ikpt < 0) return fail(c, kErrOutOfRange, "Invalid k-point index: " + std::to_string(ikpt));
    if (npw <= 0 || !K || !kinetic || !gidx)
        return fail(c, kErrRuntime, "current k-point has no G-vectors");
    // GKK::validateKineticConsistency (src/io/GKK.cppm:652-671)
    double qmax_seen = 0.0;
    for (int g = 0; g < npw; ++g) {
        double kx = K[3 * g], ky = K[3 * g + 1], kz = K[3 * g + 2];
        double e = 0.5 * (kx * kx + ky * ky + kz * kz);
        double d = kinetic[g] - e;
        if (d > 1e-10 * (1.0 + e) || d < -1e-10 * (1.0 + e))
            return fail(c, kErrRuntime, "GKK kinetic-energy mismatch at k-point " + std::to_string(ikpt) +
                                            ", G-vector " + std::to_string(g));
        qmax_seen = std::max(qmax_seen, std::sqrt(kx * kx + ky * ky + kz * kz));
    }
    // BetaqTables::interpolate throws for q > q_max (ncpp_advance.cppm:659-661)
    if (c.nch() > 0 && qmax_seen > c.tables_qmax)
        return fail(c, kErrDomain, "BetaqTables::interpolate: q exceeds q_max");
    const int n1 = c.n1, n2 = c.n2, n3 = c.n3;
    std::vector<int> xi(npw), pos(npw);
    std::vector<int> plane_of_x(n1, -1);
    std::vector<char> seen((size_t)n1 * n2 * n3, 0);
    for (int g = 0; g < npw; ++g) {
        int i = ((gidx[3 * g] % n1) + n1) % n1;
        int j = ((gidx[3 * g + 1] % n2) + n2) % n2;
        int k = ((gidx[3 * g + 2] % n3) + n3) % n3;
        size_t lin = ((size_t)i * n2 + j) * n3 + k;
        if (seen[lin])
            return fail(c, kErrInvalid, "two G-vectors map to one FFT grid point (grid too small)");
        seen[lin] = 1;
        xi[g] = i;
        pos[g] = c.hperm2[j] * n3 + c.hperm3[k];
        plane_of_x[i] = 1;
    }
    std::vector<int> plane_x;
    for (int i = 0; i < n1; ++i)
        if (plane_of_x[i] >= 0) {
            plane_of_x[i] = (int)plane_x.size();
            plane_x.push_back(i);
        }
    const int np = (int)plane_x.size();
    std::vector<int> ptr(np + 1, 0);
    for (int g = 0; g < npw; ++g) ptr[plane_of_x[xi[g]] + 1]++;
    for (int p = 0; p < np; ++p) ptr[p + 1] += ptr[p];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1), ent_ig(npw), ent_pos(npw);
    for (int g = 0; g < npw; ++g) {
        int p = plane_of_x[xi[g]];
        int e = fill[p]++;
        ent_ig[e] = g;
        ent_pos[e] = pos[g];
    }
    // planes too large for the LDS are transformed a chunk of rows at a
    // time: entries sorted by position within each plane, and the first
    // entry of every row (the order of a plane's entries does not change any
    // result: scatter positions and gather targets are distinct)
    std::vector<int> ent_rowptr;
    if (plane_lds_bytes<double>(n2, n3) > kPlaneLdsCap) {
        ent_rowptr.assign((size_t)np * (n2 + 1), 0);
        std::vector<std::pair<int, int>> pe;
        for (int p = 0; p < np; ++p) {
            pe.clear();
            for (int e = ptr[p]; e < ptr[p + 1]; ++e) pe.emplace_back(ent_pos[e], ent_ig[e]);
            std::sort(pe.begin(), pe.end());
            int* rp = ent_rowptr.data() + (size_t)p * (n2 + 1);
            for (size_t q = 0; q < pe.size(); ++q) {
                ent_pos[ptr[p] + q] = pe[q].first;
                ent_ig[ptr[p] + q] = pe[q].second;
                rp[pe[q].first / n3 + 1]++;
            }
            rp[0] = ptr[p];
            for (int r = 0; r < n2; ++r) rp[r + 1] += rp[r];
        }
    }
    // per plane: the (digit-reversed) rows that hold sphere points
    std::vector<int> row_ptr(np + 1, 0), rows;
    {
        std::vector<char> has((size_t)n2, 0);
        for (int p = 0; p < np; ++p) {
            std::fill(has.begin(), has.end(), 0);
            for (int e = ptr[p]; e < ptr[p + 1]; ++e) has[ent_pos[e] / n3] = 1;
            for (int r = 0; r < n2; ++r)
                if (has[r]) rows.push_back(r);
            row_ptr[p + 1] = (int)rows.size();
        }
    }
    if (ikpt >= (int)c.kpts.size()) c.kpts.resize(ikpt + 1);
    KptDev& k = c.kpts[ikpt];
    release_kpt(c, k);
    k.npw = npw;
    for (int d = 0; d < 3; ++d) k.kfrac[d] = kfrac ? kfrac[d] : 0.0;
    k.qmax_seen = qmax_seen;
    std::vector<double> Ks((size_t)3 * npw), kin(kinetic, kinetic + npw);
    for (int g = 0; g < npw; ++g)
        for (int d = 0; d < 3; ++d) Ks[(size_t)d * npw + g] = K[3 * g + d];
    std::vector<int> gi(gidx, gidx + 3 * (size_t)npw);
    int rc;
    if ((rc = upload(c, k.K, Ks, "K"))) return rc;
    if ((rc = upload(c, k.kin, kin, "kinetic"))) return rc;
    if ((rc = upload(c, k.gidx, gi, "gidx"))) return rc;
    if ((rc = upload(c, k.plane_x, plane_x, "plane list"))) return rc;
    if ((rc = upload(c, k.x_to_plane, plane_of_x, "plane map"))) return rc;
    if ((rc = upload(c, k.plane_ptr, ptr, "plane ptr"))) return rc;
    if ((rc = upload(c, k.ent_ig, ent_ig, "plane entries"))) return rc;
    if ((rc = upload(c, k.ent_pos, ent_pos, "plane positions"))) return rc;
    if ((rc = upload(c, k.row_ptr, row_ptr, "plane row ptr"))) return rc;
    if ((rc = upload(c, k.rows, rows, "plane rows"))) return rc;
    if ((rc = upload(c, k.ent_rowptr, ent_rowptr, "plane row entries"))) return rc;
    k.nplanes = np;
    int mx = 0;
    for (int p = 0; p < np; ++p) mx = std::max(mx, ptr[p + 1] - ptr[p]);
    k.max_plane_entries = mx;
    k.valid = true;
    return kOk;
}

int effm_kpt_dim(const effm_ctx* h, int ikpt) {
    if (!h) return kErrInvalid;
    const Ctx& c = h->c;
    if (ikpt < 0 || ikpt >= (int)c.kpts.size() || !c.kpts[ikpt].valid) return kErrOutOfRange;
    return c.kpts[ikpt].npw;
}

int effm_kpt_release(effm_ctx* h, int ikpt) {
    if (!h) return kErrInvalid;
    Ctx& c = h->c;
    if (ikpt < 0 || ikpt >= (int)c.kpts.size()) return fail(c, kErrOutOfRange, "Invalid k-point index");
    (void)hipStreamSynchronize(c.stream);
    release_kpt(c, c.kpts[ikpt]);
    return kOk;
}

static int stage_copy(Ctx& c, DevBuf& buf, const void* host, size_t bytes, bool copy, const char* what) {
    int rc = ensure(c, buf, bytes, what);
    if (rc) return rc;
    if (copy && host) HIP_TRY(hipMemcpyAsync(buf.ptr, host, bytes, hipMemcpyHostToDevice, c.stream), what);
    return kOk;
}

int effm_hpsi(effm_ctx* h, int ikpt, int nb, const void* psi, int64_t ld, void* hpsi,
              int64_t ld_out, int prec, int flags) {
    if (!h) return kErrInvalid;
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    int rc = check_kpt(c, ikpt);
    if (rc) return rc;
    KptDev& k = c.kpts[ikpt];
    if (nb < 0) return fail(c, kErrInvalid, "nb must be non-negative");
    if (nb == 0) return kOk;
    if (!psi || !hpsi) return fail(c, kErrInvalid, "psi / hpsi must not be NULL");
    if (ld < k.npw || ld_out < k.npw) return fail(c, kErrInvalid, "leading dimension smaller than npw");
    if (prec != EFFM_PREC_FP64 && prec != EFFM_PREC_MIXED) return fail(c, kErrInvalid, "unknown precision");
    if ((flags & EFFM_F_HOST) && (flags & EFFM_F_GRAPH))
        return fail(c, kErrInvalid, "EFFM_F_GRAPH takes device pointers only");
    const bool acc = !(flags & EFFM_F_OVERWRITE);
    const size_t esz = prec == EFFM_PREC_FP64 ? 16 : 8;
    const void* dpsi = psi;
    void* dh = hpsi;
    if (flags & EFFM_F_HOST) {
        if ((rc = stage_copy(c, c.stage_in, psi, esz * ld * nb, true, "staging psi"))) return rc;
        if ((rc = stage_copy(c, c.stage_out0, hpsi, esz * ld_out * nb, acc, "staging hpsi"))) return rc;
        dpsi = c.stage_in.ptr;
        dh = c.stage_out0.ptr;
    }
    ApplySeg seg{&k, dpsi, ld, {dh, nullptr, nullptr}, ld_out};
    const bool graph = (flags & EFFM_F_GRAPH) != 0;
    rc = prec == EFFM_PREC_FP64 ? apply_hpsi<double>(c, &seg, 1, nb, acc, graph, prec)
                                : apply_hpsi<float>(c, &seg, 1, nb, acc, graph, prec);
    if (rc) return rc;
    if (flags & EFFM_F_HOST) {
        HIP_TRY(hipMemcpyAsync(hpsi, dh, esz * ld_out * nb, hipMemcpyDeviceToHost, c.stream), "copy hpsi");
        HIP_TRY(hipStreamSynchronize(c.stream), "sync");
    }
    return kOk;
}

int effm_dhdk(effm_ctx* h, int ikpt, int nb, const void* psi, int64_t ld, void* ox, void* oy,
              void* oz, int64_t ld_out, int prec, int flags) {
    if (!h) return kErrInvalid;
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    int rc = check_kpt(c, ikpt);
    if (rc) return rc;
    KptDev& k = c.kpts[ikpt];
    if (nb < 0) return fail(c, kErrInvalid, "nb must be non-negative");
    if (nb == 0 || (!ox && !oy && !oz)) return kOk;
    if (!psi) return fail(c, kErrInvalid, "psi must not be NULL");
    if (ld < k.npw || ld_out < k.npw) return fail(c, kErrInvalid, "leading dimension smaller than npw");
    if (prec != EFFM_PREC_FP64 && prec != EFFM_PREC_MIXED) return fail(c, kErrInvalid, "unknown precision");
    if ((flags & EFFM_F_HOST) && (flags & EFFM_F_GRAPH))
        return fail(c, kErrInvalid, "EFFM_F_GRAPH takes device pointers only");
    const bool acc = !(flags & EFFM_F_OVERWRITE);
    const size_t esz = prec == EFFM_PREC_FP64 ? 16 : 8;
    const void* dpsi = psi;
    void* outs[3] = {ox, oy, oz};
    void* douts[3] = {ox, oy, oz};
    DevBuf* st[3] = {&c.stage_out0, &c.stage_out1, &c.stage_out2};
    if (flags & EFFM_F_HOST) {
        if ((rc = stage_copy(c, c.stage_in, psi, esz * ld * nb, true, "staging psi"))) return rc;
        dpsi = c.stage_in.ptr;
        for (int a = 0; a < 3; ++a) {
            if (!outs[a]) continue;
            if ((rc = stage_copy(c, *st[a], outs[a], esz * ld_out * nb, acc, "staging dH/dk"))) return rc;
            douts[a] = st[a]->ptr;
        }
    }
    ApplySeg seg{&k, dpsi, ld, {douts[0], douts[1], douts[2]}, ld_out};
    const bool graph = (flags & EFFM_F_GRAPH) != 0;
    rc = prec == EFFM_PREC_FP64 ? apply_dhdk<double>(c, &seg, 1, nb, acc, graph, prec)
                                : apply_dhdk<float>(c, &seg, 1, nb, acc, graph, prec);
    if (rc) return rc;
    if (flags & EFFM_F_HOST) {
        for (int a = 0; a < 3; ++a)
            if (outs[a])
                HIP_TRY(hipMemcpyAsync(outs[a], douts[a], esz * ld_out * nb, hipMemcpyDeviceToHost, c.stream),
                        "copy dH/dk");
        HIP_TRY(hipStreamSynchronize(c.stream), "sync");
    }
    return kOk;
}

// Validates a multi-k batch and builds its segments.
static int multi_segments(Ctx& c, int nk, const int* ikpts, int nb, const void* const* psi_ptrs,
                          const int64_t* ld, void* const* o0, void* const* o1, void* const* o2, int prec,
                          int flags, std::vector<ApplySeg>& segs) {
    if (nk < 0 || (nk > 0 && (!ikpts || !psi_ptrs || !ld)))
        return fail(c, kErrInvalid, "invalid batch description");
    if (flags & EFFM_F_HOST) return fail(c, kErrInvalid, "batched applies take device pointers only");
    if (nb < 0) return fail(c, kErrInvalid, "nb must be non-negative");
    if (prec != EFFM_PREC_FP64 && prec != EFFM_PREC_MIXED) return fail(c, kErrInvalid, "unknown precision");
    segs.clear();
    for (int i = 0; i < nk; ++i) {
        int rc = check_kpt(c, ikpts[i]);
        if (rc) return rc;
        const KptDev& k = c.kpts[ikpts[i]];
        void* out[3] = {o0 ? o0[i] : nullptr, o1 ? o1[i] : nullptr, o2 ? o2[i] : nullptr};
        if (!psi_ptrs[i]) return fail(c, kErrInvalid, "psi must not be NULL");
        if (ld[i] < k.npw) return fail(c, kErrInvalid, "leading dimension smaller than npw");
        segs.push_back(ApplySeg{&k, psi_ptrs[i], ld[i], {out[0], out[1], out[2]}, ld[i]});
    }
    return kOk;
}

int effm_hpsi_multi(effm_ctx* h, int nk, const int* ikpts, int nb, const void* const* psi_ptrs,
                    const int64_t* ld, void* const* hpsi_ptrs, int prec, int flags) {
    if (!h) return kErrInvalid;
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    std::vector<ApplySeg> segs;
    int rc = multi_segments(c, nk, ikpts, nb, psi_ptrs, ld, hpsi_ptrs, nullptr, nullptr, prec, flags, segs);
    if (rc) return rc;
    if (nk > 0 && !hpsi_ptrs) return fail(c, kErrInvalid, "hpsi pointers must not be NULL");
    for (const auto& s : segs)
        if (!s.out[0]) return fail(c, kErrInvalid, "hpsi must not be NULL");
    if (nk == 0 || nb == 0) return kOk;
    const bool acc = !(flags & EFFM_F_OVERWRITE);
    const bool graph = (flags & EFFM_F_GRAPH) != 0;
    return prec == EFFM_PREC_FP64 ? apply_hpsi<double>(c, segs.data(), nk, nb, acc, graph, prec)
                                  : apply_hpsi<float>(c, segs.data(), nk, nb, acc, graph, prec);
}

int effm_dhdk_multi(effm_ctx* h, int nk, const int* ikpts, int nb, const void* const* psi_ptrs,
                    const int64_t* ld, void* const* ox, void* const* oy, void* const* oz, int prec, int flags) {
    if (!h) return kErrInvalid;
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    std::vector<ApplySeg> segs;
    int rc = multi_segments(c, nk, ikpts, nb, psi_ptrs, ld, ox, oy, oz, prec, flags, segs);
    if (rc) return rc;
    if (nk == 0 || nb == 0 || (!ox && !oy && !oz)) return kOk;
    const bool acc = !(flags & EFFM_F_OVERWRITE);
    const bool graph = (flags & EFFM_F_GRAPH) != 0;
    return prec == EFFM_PREC_FP64 ? apply_dhdk<double>(c, segs.data(), nk, nb, acc, graph, prec)
                                  : apply_dhdk<float>(c, segs.data(), nk, nb, acc, graph, prec);
}

int effm_hpsi_dhdk_multi(effm_ctx* h, int nk, const int* ikpts, int nb, const void* const* psi_ptrs,
                         const int64_t* ld, void* const* hpsi_ptrs, void* const* ox, void* const* oy,
                         void* const* oz, int prec, int flags) {
    if (!h) return kErrInvalid;
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    std::vector<ApplySeg> segs;
    int rc = multi_segments(c, nk, ikpts, nb, psi_ptrs, ld, ox, oy, oz, prec, flags, segs);
    if (rc) return rc;
    if (nk > 0 && !hpsi_ptrs) return fail(c, kErrInvalid, "hpsi pointers must not be NULL");
    for (int i = 0; i < nk; ++i) {
        if (!hpsi_ptrs[i]) return fail(c, kErrInvalid, "hpsi must not be NULL");
        segs[i].hout = hpsi_ptrs[i];
    }
    if (nk == 0 || nb == 0) return kOk;
    const bool acc = !(flags & EFFM_F_OVERWRITE);
    const bool graph = (flags & EFFM_F_GRAPH) != 0;
    return prec == EFFM_PREC_FP64 ? apply_hd<double>(c, segs.data(), nk, nb, acc, graph, prec)
                                  : apply_hd<float>(c, segs.data(), nk, nb, acc, graph, prec);
}

int effm_autotune(effm_ctx* h, int ikpt, int nb, int prec, int* pencil_bands, int* gemm_wgs) {
    if (!h) return kErrInvalid;
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    int rc = check_kpt(c, ikpt);
    if (rc) return rc;
    if (nb <= 0) return fail(c, kErrInvalid, "nb must be positive");
    if (prec != EFFM_PREC_FP64 && prec != EFFM_PREC_MIXED) return fail(c, kErrInvalid, "unknown precision");
    KptDev& k = c.kpts[ikpt];
    return prec == EFFM_PREC_FP64 ? autotune<double>(c, k, nb, pencil_bands, gemm_wgs)
                                  : autotune<float>(c, k, nb, pencil_bands, gemm_wgs);
}

static bool one_of(int v, const int* choices, int n) {
    for (int i = 0; i < n; ++i)
        if (choices[i] == v) return true;
    return false;
}

int effm_get_shapes(const effm_ctx* h, int prec, effm_shapes* out) {
    if (!h || !out) return kErrInvalid;
    if (prec != EFFM_PREC_FP64 && prec != EFFM_PREC_MIXED) return kErrInvalid;
    const Ctx& c = h->c;
    const Ctx::Tune& tn = c.tune[prec == EFFM_PREC_FP64 ? 0 : 1];
    out->pencil_bands = tn.pencil_bands;
    out->pencil_z = prec == EFFM_PREC_FP64 ? pencil_z_in_effect<double>(c) : pencil_z_in_effect<float>(c);
    out->gemm_wgs = tn.gemm_wgs;
    out->apply_tile = tn.apply_tile;
    return kOk;
}

int effm_set_shapes(effm_ctx* h, int prec, const effm_shapes* in) {
    if (!h) return kErrInvalid;
    Ctx& c = h->c;
    if (!in) return fail(c, kErrInvalid, "shapes is NULL");
    if (prec != EFFM_PREC_FP64 && prec != EFFM_PREC_MIXED) return fail(c, kErrInvalid, "unknown precision");
    if (!one_of(in->pencil_bands, kPencilBandChoices, (int)std::size(kPencilBandChoices)))
        return fail(c, kErrInvalid, "pencil_bands must be 4, 8 or 16");
    if (in->pencil_z != 0 && !one_of(in->pencil_z, kPencilZChoices, (int)std::size(kPencilZChoices)))
        return fail(c, kErrInvalid, "pencil_z must be 0, 16 or 18");
    if (!one_of(in->gemm_wgs, kGemmWgChoices, (int)std::size(kGemmWgChoices)))
        return fail(c, kErrInvalid, "gemm_wgs must be 512, 768, 1024, 1536 or 2048");
    if (!one_of(in->apply_tile, kApplyTileChoices, (int)std::size(kApplyTileChoices)))
        return fail(c, kErrInvalid, "apply_tile must be 64 or 32");
    Ctx::Tune& tn = c.tune[prec == EFFM_PREC_FP64 ? 0 : 1];
    tn.pencil_bands = in->pencil_bands;
    tn.pencil_z = in->pencil_z;
    tn.gemm_wgs = in->gemm_wgs;
    tn.apply_tile = in->apply_tile;
    ++c.epoch;    // captured graphs hold the old shapes: recapture
    return kOk;
}

int effm_overlap(effm_ctx* h, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb,
                 void* C, int64_t ldc, int prec, int flags) {
    if (!h) return kErrInvalid;
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    if (m < 0 || n < 0 || k < 0) return fail(c, kErrInvalid, "negative overlap dimension");
    if (m == 0 || n == 0) return kOk;
    if (!C || (k > 0 && (!A || !B))) return fail(c, kErrInvalid, "overlap operands must not be NULL");
    if (k > 0 && (lda < k || ldb < k)) return fail(c, kErrInvalid, "leading dimension smaller than k");
    if (ldc < n) return fail(c, kErrInvalid, "ldc smaller than n");
    if (prec != EFFM_PREC_FP64 && prec != EFFM_PREC_MIXED) return fail(c, kErrInvalid, "unknown precision");
    if (flags & (EFFM_F_HOST | EFFM_F_GRAPH)) return fail(c, kErrInvalid, "effm_overlap takes device pointers only");
    const size_t pb = prec == EFFM_PREC_FP64 ? overlap_partial_bytes<double>(c, m, n, std::max(k, 1))
                                             : overlap_partial_bytes<float>(c, m, n, std::max(k, 1));
    int rc = ensure(c, c.partial, pb, "overlap partial sums");
    if (rc) return rc;
    hipError_t e = prec == EFFM_PREC_FP64
                       ? launch_overlap<double>(c, m, n, k, A, lda, B, ldb, C, ldc, c.partial.ptr, c.partial.bytes)
                       : launch_overlap<float>(c, m, n, k, A, lda, B, ldb, C, ldc, c.partial.ptr, c.partial.bytes);
    if (e != hipSuccess) return hip_fail(c, e, "overlap");
    return kOk;
}

int effm_overlap_noalloc(effm_ctx* h, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb,
                         void* C, int64_t ldc, int prec, int flags) {
    if (!h) return kErrInvalid;
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    if (m < 0 || n < 0 || k < 0) return fail(c, kErrInvalid, "negative overlap dimension");
    if (m == 0 || n == 0) return kOk;
    if (!C || (k > 0 && (!A || !B))) return fail(c, kErrInvalid, "overlap operands must not be NULL");
    if (k > 0 && (lda < k || ldb < k)) return fail(c, kErrInvalid, "leading dimension smaller than k");
    if (ldc < n) return fail(c, kErrInvalid, "ldc smaller than n");
    if (prec != EFFM_PREC_FP64 && prec != EFFM_PREC_MIXED) return fail(c, kErrInvalid, "unknown precision");
    if (flags & (EFFM_F_HOST | EFFM_F_GRAPH))
        return fail(c, kErrInvalid, "effm_overlap_noalloc takes device pointers only");
    // the same split count as effm_overlap, so the same partial sums
    const size_t pb = prec == EFFM_PREC_FP64 ? overlap_partial_bytes<double>(c, m, n, std::max(k, 1))
                                             : overlap_partial_bytes<float>(c, m, n, std::max(k, 1));
    if (pb > kOverlapPartialCap)
        return fail(c, kErrInvalid, "effm_overlap_noalloc: an m x n block of partial sums (" + std::to_string(pb) +
                                        " B) exceeds the " + std::to_string(kOverlapPartialCap >> 20) +
                                        " MiB static scratch");
    void* part = scratch_ptr(c);
    if (!part) return kErrHip;
    // every partial sum is written by the GEMM before the reduction reads it,
    // so the scratch needs no clearing between calls
    hipError_t e = prec == EFFM_PREC_FP64
                       ? launch_overlap<double>(c, m, n, k, A, lda, B, ldb, C, ldc, part, kOverlapPartialCap, false)
                       : launch_overlap<float>(c, m, n, k, A, lda, B, ldb, C, ldc, part, kOverlapPartialCap, false);
    if (e != hipSuccess) return hip_fail(c, e, "overlap (no allocation)");
    return kOk;
}

int effm_lincomb(effm_ctx* h, int64_t n, double a, void* z, double b, const void* y, double cc, const void* x,
                 int prec, int flags) {
    if (!h) return kErrInvalid;
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    if (n < 0) return fail(c, kErrInvalid, "negative length");
    if (n == 0) return kOk;
    if (!z || !y || !x) return fail(c, kErrInvalid, "lincomb operands must not be NULL");
    if (prec != EFFM_PREC_FP64 && prec != EFFM_PREC_MIXED) return fail(c, kErrInvalid, "unknown precision");
    if (flags & (EFFM_F_HOST | EFFM_F_GRAPH)) return fail(c, kErrInvalid, "effm_lincomb takes device pointers only");
    const hipError_t e = prec == EFFM_PREC_FP64 ? launch_lincomb<double>(c, n, a, z, b, y, cc, x)
                                                : launch_lincomb<float>(c, n, a, z, b, y, cc, x);
    if (e != hipSuccess) return hip_fail(c, e, "lincomb");
    return kOk;
}

int effm_rayleigh(effm_ctx* h, int npw, int nb, const void* psi, int64_t ld, const void* hpsi,
                  int64_t ld_h, int prec, double* E, int flags) {
    if (!h) return kErrInvalid;
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    if (prec != EFFM_PREC_FP64 && prec != EFFM_PREC_MIXED) return fail(c, kErrInvalid, "unknown precision");
    if (nb <= 0) return kOk;
    if (!psi || !hpsi || !E || npw <= 0 || ld < npw || ld_h < npw)
        return fail(c, kErrInvalid, "invalid Rayleigh-quotient arguments");
    const size_t esz = prec == EFFM_PREC_FP64 ? 16 : 8;
    const void* dp = psi;
    const void* dh = hpsi;
    double* dE = E;
    int rc;
    if (flags & EFFM_F_HOST) {
        if ((rc = stage_copy(c, c.stage_in, psi, esz * ld * nb, true, "staging psi"))) return rc;
        if ((rc = stage_copy(c, c.stage_out0, hpsi, esz * ld_h * nb, true, "staging hpsi"))) return rc;
        if ((rc = ensure(c, c.stage_out1, sizeof(double) * nb, "staging E"))) return rc;
        dp = c.stage_in.ptr;
        dh = c.stage_out0.ptr;
        dE = static_cast<double*>(c.stage_out1.ptr);
    }
    timer_begin(c);
    hipError_t e = prec == EFFM_PREC_FP64 ? launch_rayleigh<double>(c.stream, npw, nb, dp, ld, dh, ld_h, dE)
                                          : launch_rayleigh<float>(c.stream, npw, nb, dp, ld, dh, ld_h, dE);
    timer_end(c, "rayleigh");
    if (e != hipSuccess) return hip_fail(c, e, "rayleigh");
    if (flags & EFFM_F_HOST) {
        HIP_TRY(hipMemcpyAsync(E, dE, sizeof(double) * nb, hipMemcpyDeviceToHost, c.stream), "copy E");
        HIP_TRY(hipStreamSynchronize(c.stream), "sync");
    }
    return kOk;
}

int effm_kp_mass(effm_ctx* h, int nb, const double* E, const void* const* V, int64_t ldv, int nsel,
                 const int* bands, const double* d2h, double degen_tol, double* inv_mass, int flags) {
    if (!h) return kErrInvalid;
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    if (nsel < 0 || nb < 0) return fail(c, kErrInvalid, "negative band count");
    if (nsel == 0) return kOk;
    if (nsel > kKpMaxBands)
        return fail(c, kErrInvalid, "at most " + std::to_string(kKpMaxBands) + " bands per effm_kp_mass call");
    if (!E || !V || !V[0] || !V[1] || !V[2] || !bands || !d2h || !inv_mass)
        return fail(c, kErrInvalid, "k.p operands must not be NULL");
    if (ldv < nb) return fail(c, kErrInvalid, "ldv smaller than nb");
    if (!(degen_tol >= 0.0)) return fail(c, kErrInvalid, "degeneracy tolerance must be non-negative");
    if (flags & (EFFM_F_HOST | EFFM_F_GRAPH)) return fail(c, kErrInvalid, "effm_kp_mass takes device pointers only");
    KpBands bs{};
    for (int i = 0; i < nsel; ++i) {
        if (bands[i] < 0 || bands[i] >= nb) return fail(c, kErrOutOfRange, "band index out of range");
        bs.n[i] = bands[i];
    }
    timer_begin(c);
    hipError_t e = launch_kp_sum(c.stream, bs, nsel, nb, E, V, ldv, d2h, degen_tol, inv_mass);
    timer_end(c, "kp_sum");
    if (e != hipSuccess) return hip_fail(c, e, "k.p sum");
    return kOk;
}

int effm_set_timing(effm_ctx* h, int enable) {
    if (!h) return kErrInvalid;
    h->c.timing = enable != 0;
    return kOk;
}

int effm_get_timing(const effm_ctx* hc, const char* name, double* ms, int64_t* count) {
    if (!hc || !name) return kErrInvalid;
    effm_ctx* h = const_cast<effm_ctx*>(hc);
    timer_resolve(h->c);
    auto it = h->c.timers.find(name);
    if (it == h->c.timers.end()) {
        if (ms) *ms = 0.0;
        if (count) *count = 0;
        return kOk;
    }
    if (ms) *ms = it->second.ms;
    if (count) *count = it->second.count;
    return kOk;
}

int effm_reset_timing(effm_ctx* h) {
    if (!h) return kErrInvalid;
    timer_resolve(h->c);
    h->c.timers.clear();
    return kOk;
}

}  // extern "C"
