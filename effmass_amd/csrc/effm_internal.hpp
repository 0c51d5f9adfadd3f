// effm_internal.hpp — context layout and kernel launchers of libeffm_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "fft_lds.hpp"

namespace effm {

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup i runs on
// XCD i % 8), and each XCD has its own L2.  xcd_order(i, G) renumbers a 1D
// grid of G workgroups so that consecutive logical ids run on one XCD:
// workgroups that read the same lines (GEMM tiles on the same operand rows,
// neighbouring z-chunks of the pencil pass) are given consecutive ids and so
// share one L2.  A bijection for any G (the remainder keeps its ids).
constexpr int kXcds = 8;
__device__ inline int xcd_order(int i, int G) {
    const int per = G / kXcds;
    if (i >= per * kXcds) return i;
    return (i % kXcds) * per + i / kXcds;
}

// Status codes (mirrors include/effm_hpsi.h).
enum : int {
    kOk = 0,
    kErrRuntime = -1,
    kErrOutOfRange = -2,
    kErrDomain = -3,
    kErrInvalid = -4,
    kErrNoMem = -5,
    kErrHip = -6,
};

struct Status {
    int code = kOk;
    std::string msg;
    static Status ok() { return {}; }
    bool good() const { return code == kOk; }
};

// Device buffer owned by the context (counted against the HBM budget).
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// Projector channel: one (atom, l, ib, m) with |lambda| >= 1e-15, in the
// reference's loop order type -> atom -> (l, ib) -> m
// (src/H_psi/hamiltonian_callable.cpp:211-261).
struct Channel {
    int atom;
    int table;      // row of the concatenated beta / dbeta tables
    int l;
    int m;
    double lambda;
    int species;
};

struct TableInfo {   // per table row
    int nq;
    double dq;
    double qmax;
    double dbeta0;   // dbeta/dq(0) * 1/sqrt(Omega), the q->0 limit of beta/q
};

struct KptDev {
    bool valid = false;
    int npw = 0;
    double kfrac[3] = {0, 0, 0};
    double qmax_seen = 0;
    // SoA K (3*npw), kinetic (npw), gidx (3*npw)
    DevBuf K, kin, gidx;
    // plane CSR: active x-planes, entries grouped by plane
    int nplanes = 0;
    DevBuf plane_x;      // int[nplanes]
    DevBuf x_to_plane;   // int[n1], -1 if inactive
    DevBuf plane_ptr;    // int[nplanes+1]
    DevBuf ent_ig;       // int[npw]
    DevBuf ent_pos;      // int[npw] digit-reversed (j,k) position in the plane
    DevBuf row_ptr;      // int[nplanes+1]: per plane, its rows holding sphere points
    DevBuf rows;         //   (digit-reversed row index j', ascending)
    int max_plane_entries = 0;
    // grids whose n2 x n3 FP64 plane exceeds the LDS: entries sorted by
    // position within each plane, and per plane the first entry of each
    // (digit-reversed) row, [nplanes][n2 + 1]
    DevBuf ent_rowptr;
    // projector matrices, built on first use: P[c][g]; dP[alpha][c][g]
    DevBuf P64, P32, dP64, dP32;
};

struct Timer {
    double ms = 0.0;
    int64_t count = 0;
};

struct Ctx {
    int device = 0;
    int ncu = 256;                      // compute units of the device
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    uint64_t used = 0;
    std::string err;

    // cell
    bool have_cell = false;
    int n1 = 0, n2 = 0, n3 = 0;
    double A[9] = {0};
    double omega = 0;
    int natom = 0;
    int lmax = -1;
    FftAxis ax1{}, ax2{}, ax3{};
    DevBuf vloc64, vloc32;
    DevBuf tw1_64, tw2_64, tw3_64, tw1_32, tw2_32, tw3_32;
    DevBuf perm1;                       // int[n1]
    std::vector<int> hperm1, hperm2, hperm3;
    DevBuf tau;                         // double[natom*3]
    std::vector<Channel> chans;
    std::vector<TableInfo> tinfo;
    double tables_qmax = 0;             // min qmax over tables
    DevBuf btab, dbtab;                 // double[ntab*nq]
    DevBuf tab_info;                    // TableDev[ntab]
    DevBuf chan_dev;                    // ChanDev[nch]
    DevBuf lam64;                       // double[nch]
    bool have_dbeta = false;
    int nq_max = 0;

    std::vector<KptDev> kpts;

    // per kernel: resident workgroups per CU (dynamic LDS attribute set once)
    std::map<const void*, int> occupancy;

    // launch shapes, per precision (0 = FP64, 1 = mixed): bands and z
    // points per x-pencil workgroup (z 0 = the grid's default), the
    // workgroup count the split-K projection aims for and the band-tile
    // width of the apply GEMMs above 32 bands.  Defaults are the measured
    // best on the 90^3 / 64-atom cell; effm_autotune re-picks them by timing
    // the candidates on this cell.
    struct Tune {
        int pencil_bands = 8;
        int pencil_z = 0;
        int gemm_wgs = 1024;
        int apply_tile = 64;
        bool tuned = false;
    };
    Tune tune[2];
    template <typename T>
    const Tune& tuning() const { return tune[sizeof(T) == 8 ? 0 : 1]; }

    // captured apply sequences (EFFM_F_GRAPH), keyed by everything the
    // launches read; any device (re)allocation bumps epoch and drops them
    struct Graph {
        std::vector<int64_t> key;
        hipGraphExec_t exec;
        uint64_t uses;
    };
    std::vector<Graph> graphs;
    uint64_t epoch = 0;
    uint64_t graph_epoch = 0;
    uint64_t graph_clock = 0;
    hipStream_t capture_stream = nullptr;
    bool capturing = false;

    // scratch (grown on demand)
    DevBuf work;       // FFT plane workspace
    DevBuf partial;    // split-K partial sums
    DevBuf cbuf;       // reduced <p|psi> (scaled), precision of the path
    DevBuf stage_in, stage_out0, stage_out1, stage_out2;   // host-pointer staging

    bool timing = false;
    std::map<std::string, Timer> timers;
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    hipEvent_t cur_start = nullptr;

    int nch() const { return static_cast<int>(chans.size()); }
};

// ---- device-side descriptors -------------------------------------------
struct ChanDev {
    int atom;
    int table;
    int l;
    int m;
};

struct TableDev {
    int nq;
    int pad;
    double dq;
    double qmax;
    double norm;     // 1/sqrt(Omega)
    double dbeta0;   // normalised dbeta/dq at q = 0
};

// Largest dynamic LDS of one workgroup on gfx950; planes (plus twiddles)
// beyond it take the split-plane kernels.
constexpr size_t kPlaneLdsCap = 160 * 1024;

// LDS bytes of one whole-plane workgroup (plane + both axes' twiddles).
template <typename T>
inline size_t plane_lds_bytes(int n2, int n3) {
    return sizeof(typename Cx<T>::type) * ((size_t)n2 * n3 + n2 + n3);
}

// Row pitch (complex elements) of the FFT plane workspace.  FP32 rows are
// padded to a 128-byte line (the x-pencil pass then reads whole lines);
// FP64 rows stay unpadded (measured faster on the 90^3 grid: the plane
// passes are limited by their own LDS work, not by line alignment).
template <typename T>
constexpr int ws_pitch(int n3) {
    constexpr int q = 128 / (int)(2 * sizeof(T));
    return sizeof(T) == 8 ? n3 : (n3 + q - 1) / q * q;
}

// Plane stride (complex elements) of the workspace.
template <typename T>
constexpr int64_t ws_plane(int n2, int n3) {
    return (int64_t)n2 * ws_pitch<T>(n3);
}

// ---- batched applies ------------------------------------------------------
// One launch of each kernel covers up to kMaxSeg k-points ("segments") with
// their own ragged G-sphere, plane CSR, projector matrices and psi / output
// buffers; the per-segment descriptors travel as kernel arguments.
constexpr int kMaxSeg = 16;

// Host-side description of one k-point of a batched apply.
struct ApplySeg {
    const KptDev* k;
    const void* psi;
    int64_t ld;
    void* out[3];         // H|psi>: out[0]; dH/dk: x, y, z (NULL = skip)
    int64_t ldo;
    void* hout = nullptr; // dH/dk together with H|psi>: the H|psi> output
};

// ---- launchers (k_local.hip) --------------------------------------------
// bands [b0, b0 + nb) of every segment; work holds the segments' planes back
// to back (local_work_bytes)
template <typename T>
hipError_t launch_local(Ctx& c, const ApplySeg* segs, int nseg, int b0, int nb, bool accumulate,
                        void* work);

template <typename T>
size_t local_work_bytes(const Ctx& c, const ApplySeg* segs, int nseg, int nb);

// True if the compile-time square-plane kernels serve this cell with
// coprime radices p1 x p2 = n2 (then in prime-factor order, see fft_ct.hpp).
bool plane_ct_pfa(const Ctx& c, int* p1, int* p2);

// ---- launchers (k_nonlocal.hip) -----------------------------------------
template <typename T>
hipError_t launch_build_proj(Ctx& c, KptDev& k, bool grad);

template <typename T>
hipError_t launch_vnl(Ctx& c, const ApplySeg* segs, int nseg, int nb, void* partial,
                      size_t partial_bytes, void* cbuf);

// with_h: also the nonlocal part of H|psi> into each segment's hout (added
// to what is there), from the <p|psi> block of the same projection
template <typename T>
hipError_t launch_dhdk(Ctx& c, const ApplySeg* segs, int nseg, int nb, bool accumulate,
                       void* partial, size_t partial_bytes, void* cbuf,
                       bool with_h = false);

// split-K partial sums of nseg segments with nrows projector rows each
template <typename T>
size_t vnl_partial_bytes(Ctx& c, int nrows, int nb, int min_npw, int nseg);

// overlap C = conj(A) B^T with FP64-reduced split-K partial sums; the split
// count keeps the partial sums within kOverlapPartialCap bytes whenever one
// split's m*n block fits (an eigensolver's context runs under a small budget)
constexpr size_t kOverlapPartialCap = size_t(32) << 20;
template <typename T>
size_t overlap_partial_bytes(Ctx& c, int m, int n, int k);
template <typename T>
hipError_t launch_overlap(Ctx& c, int m, int n, int k, const void* A, int64_t lda, const void* B, int64_t ldb,
                          void* C, int64_t ldc, void* partial, size_t partial_bytes, bool timed = true);
void prime_occupancy(Ctx& c);

// candidate launch shapes effm_autotune times
constexpr int kPencilBandChoices[] = {4, 8, 16};
constexpr int kPencilZChoices[] = {16, 18};
constexpr int kGemmWgChoices[] = {512, 768, 1024, 1536, 2048};
constexpr int kApplyTileChoices[] = {64, 32};
// z points per x-pencil workgroup in effect for precision T on this grid
template <typename T>
int pencil_z_in_effect(const Ctx& c);

template <typename T>
hipError_t launch_lincomb(Ctx& c, int64_t n, double a, void* z, double b, const void* y, double cc,
                          const void* x);

// selected bands of one k.p launch (by value)
constexpr int kKpMaxBands = 64;
struct KpBands {
    int n[kKpMaxBands];
};
hipError_t launch_kp_sum(hipStream_t s, const KpBands& bs, int nsel, int nb, const double* E,
                         const void* const* V, int64_t ldv, const double* d2h, double tol, double* out);

template <typename T>
hipError_t launch_rayleigh(hipStream_t s, int npw, int nb, const void* psi, int64_t ld,
                           const void* hpsi, int64_t ldh, double* E);

// timing helpers (effm_api.hip)
void timer_begin(Ctx& c);
void timer_end(Ctx& c, const char* name);

}  // namespace effm
