// k_local.hip — kinetic + local-potential part of H|psi> on gfx950.
//
// Reference: src/H_psi/hamiltonian_callable.cpp:119-182 (kinetic, scatter,
// G2R FFT, V_loc multiply, R2G FFT, gather) and src/math/fft.cppm:60-85.
//
// Three launches per band batch, each an HBM streaming pass:
//   1. plane_g2r     one item per (active x-plane, band): scatter the
//                    plane's G-sphere entries into a zeroed n2 x n3 LDS
//                    plane at digit-reversed positions, DIT backward FFT
//                    along k then j, write the plane.
//   2. xpencil_vloc  one workgroup per (j, z-chunk, band group): load the
//                    x-pencils of the active planes (inactive x = 0),
//                    backward FFT along x, multiply by V_loc(r), forward FFT
//                    along x, store only the active planes (the R2G result
//                    is needed only inside the sphere).
//   3. plane_r2g     one item per (active x-plane, band): forward FFT along
//                    j and k (DIF), then gather the sphere entries:
//                    hpsi = [hpsi +] kin*psi + V psi / N.
// Planes outside the sphere's x-range are never written or read: this is
// the zero-padding the reference's dense 3D FFT spends on them.
//
// Every launch covers a batch of up to kMaxSeg k-points ("segments"), each
// with its own ragged G-sphere (npw), plane CSR, psi / hpsi and workspace;
// the segment descriptors are kernel arguments.  The plane kernels walk the
// concatenated (segment, plane, band) item list; the pencil kernel's grid
// is (j, z-chunk) x (segment, band group).
#include <algorithm>

#include "effm_internal.hpp"
#include "fft_ct.hpp"

namespace effm {

template <typename T>
struct LocalSeg {
    const cx_t<T>* psi;       // band b0 of the batch
    cx_t<T>* hpsi;
    int64_t ld, ldo;
    const int* plane_ptr;     // [nplanes + 1]
    const int* ent_ig;        // [npw]
    const int* ent_pos;       // [npw]
    const int* x_to_plane;    // [n1]
    const int* plane_x;       // [nplanes]
    const int* row_ptr;       // [nplanes + 1]
    const int* rows;          // rows of each plane that hold sphere points
    const double* kin;        // [npw]
    const int* ent_rowptr;    // [nplanes][n2 + 1] (split-plane grids) or NULL
    cx_t<T>* work;            // [nb][nplanes][ws_plane] or [nb][n2][nplanes][pitch]
    // element (band b, plane p, row j, z) of the compile-time kernels sits
    // at work[b nplanes ws_plane + p wsp + j wsr + z]: plane-major
    // (wsp = ws_plane, wsr = pitch) or row-major (wsp = pitch,
    // wsr = nplanes pitch: the x-pencil pass of one row j reads one
    // contiguous block of all planes); the runtime-plan kernels use
    // plane-major
    int64_t wsp, wsr;
    int nplanes;
    int item0;                // first (plane, band) item of the segment
};

template <typename T>
struct LocalBatch {
    int nseg;
    int nb;                   // bands per segment
    LocalSeg<T> s[kMaxSeg];
};

// Segment holding flat item `item` (segments are in ascending item0).
template <typename T>
__device__ __forceinline__ int seg_of(const LocalBatch<T>& bt, int item) {
    int s = 0;
    while (s + 1 < bt.nseg && item >= bt.s[s + 1].item0) ++s;
    return s;
}

// ---------------------------------------------------------------------------
//  Generic kernels (any n1, n2, n3 the runtime radix set factors).
//  Grids: planes (max nplanes, nb, nseg); pencils (n2 * z-chunks, nb, nseg).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(512)
k_plane_g2r(const LocalBatch<T> bt, FftAxis a2, FftAxis a3, const cx_t<T>* __restrict__ tw2,
            const cx_t<T>* __restrict__ tw3) {
    using C = cx_t<T>;
    const LocalSeg<T>& sg = bt.s[blockIdx.z];
    const int p = blockIdx.x;
    if (p >= sg.nplanes) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n2 = a2.n, n3 = a3.n, n23 = n2 * n3;
    const int wp = ws_pitch<T>(n3);
    C* plane = reinterpret_cast<C*>(smem_raw);
    C* t2 = plane + n23;
    C* t3 = t2 + n2;
    const int b = blockIdx.y;
    for (int i = threadIdx.x; i < n23; i += blockDim.x) plane[i] = C{0, 0};
    for (int i = threadIdx.x; i < n2; i += blockDim.x) t2[i] = tw2[i];
    for (int i = threadIdx.x; i < n3; i += blockDim.x) t3[i] = tw3[i];
    __syncthreads();
    const C* ps = sg.psi + (int64_t)b * sg.ld;
    const int e0 = sg.plane_ptr[p], e1 = sg.plane_ptr[p + 1];
    for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) plane[sg.ent_pos[e]] = ps[sg.ent_ig[e]];
    __syncthreads();
    fft_lds<+1, true, T>(plane, a3, 1, n3, n2, t3);    // along k (rows)
    fft_lds<+1, true, T>(plane, a2, n3, 1, n3, t2);    // along j (columns)
    C* out = sg.work + ((int64_t)b * sg.nplanes + p) * ws_plane<T>(n2, n3);
    for (int i = threadIdx.x; i < n23; i += blockDim.x) {
        int j = i / n3;
        out[j * wp + (i - j * n3)] = plane[i];
    }
}

template <typename T, int TZ>
__global__ void __launch_bounds__(256)
k_xpencil_vloc(const LocalBatch<T> bt, const int* __restrict__ perm1, const T* __restrict__ vloc,
               FftAxis a1, int n2, int n3, const cx_t<T>* __restrict__ tw1) {
    using C = cx_t<T>;
    const LocalSeg<T>& sg = bt.s[blockIdx.z];
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n1 = a1.n;
    C* buf = reinterpret_cast<C*>(smem_raw);      // [n1][TZ]
    C* t1 = buf + n1 * TZ;
    const int nzc = (n3 + TZ - 1) / TZ;
    const int j = blockIdx.x / nzc;
    const int z0 = (blockIdx.x - j * nzc) * TZ;
    const int nz = min(TZ, n3 - z0);
    const int b = blockIdx.y;
    const int nplanes = sg.nplanes;
    const int64_t n23 = (int64_t)n2 * n3;
    const int64_t wps = ws_plane<T>(n2, n3);   // workspace plane stride
    cx_t<T>* wb = sg.work + (int64_t)b * nplanes * wps + (int64_t)j * ws_pitch<T>(n3) + z0;
    for (int i = threadIdx.x; i < n1; i += blockDim.x) t1[i] = tw1[i];
    for (int idx = threadIdx.x; idx < n1 * TZ; idx += blockDim.x) {
        int x = idx / TZ, z = idx - x * TZ;
        int pl = sg.x_to_plane[x];
        C v = C{0, 0};
        if (pl >= 0 && z < nz) v = wb[(int64_t)pl * wps + z];
        buf[perm1[x] * TZ + z] = v;
    }
    __syncthreads();
    fft_lds<+1, true, T>(buf, a1, TZ, 1, TZ, t1);      // G2R along x
    const T* vb = vloc + (int64_t)j * n3 + z0;
    for (int idx = threadIdx.x; idx < n1 * TZ; idx += blockDim.x) {
        int x = idx / TZ, z = idx - x * TZ;
        if (z < nz) {
            T v = vb[(int64_t)x * n23 + z];
            C a = buf[idx];
            buf[idx] = C{a.x * v, a.y * v};
        }
    }
    __syncthreads();
    fft_lds<-1, false, T>(buf, a1, TZ, 1, TZ, t1);     // R2G along x
    for (int idx = threadIdx.x; idx < nplanes * TZ; idx += blockDim.x) {
        int pl = idx / TZ, z = idx - pl * TZ;
        if (z < nz) wb[(int64_t)pl * wps + z] = buf[perm1[sg.plane_x[pl]] * TZ + z];
    }
}

template <typename T>
__global__ void __launch_bounds__(512)
k_plane_r2g(const LocalBatch<T> bt, T scale, int accumulate, FftAxis a2, FftAxis a3,
            const cx_t<T>* __restrict__ tw2, const cx_t<T>* __restrict__ tw3) {
    using C = cx_t<T>;
    const LocalSeg<T>& sg = bt.s[blockIdx.z];
    const int p = blockIdx.x;
    if (p >= sg.nplanes) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n2 = a2.n, n3 = a3.n, n23 = n2 * n3;
    const int wp = ws_pitch<T>(n3);
    C* plane = reinterpret_cast<C*>(smem_raw);
    C* t2 = plane + n23;
    C* t3 = t2 + n2;
    const int b = blockIdx.y;
    const C* in = sg.work + ((int64_t)b * sg.nplanes + p) * ws_plane<T>(n2, n3);
    for (int i = threadIdx.x; i < n23; i += blockDim.x) {
        int j = i / n3;
        plane[i] = in[j * wp + (i - j * n3)];
    }
    for (int i = threadIdx.x; i < n2; i += blockDim.x) t2[i] = tw2[i];
    for (int i = threadIdx.x; i < n3; i += blockDim.x) t3[i] = tw3[i];
    __syncthreads();
    fft_lds<-1, false, T>(plane, a2, n3, 1, n3, t2);   // along j
    fft_lds<-1, false, T>(plane, a3, 1, n3, n2, t3);   // along k
    const C* ps = sg.psi + (int64_t)b * sg.ld;
    C* hs = sg.hpsi + (int64_t)b * sg.ldo;
    const int e0 = sg.plane_ptr[p], e1 = sg.plane_ptr[p + 1];
    for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        int ig = sg.ent_ig[e];
        C v = plane[sg.ent_pos[e]];
        C s = ps[ig];
        double kn = sg.kin[ig];
        // kinetic term in FP64 (callable.cpp:119-126), then the gathered
        // local-potential term (callable.cpp:171-182)
        double re = kn * (double)s.x + (double)(v.x * scale);
        double im = kn * (double)s.y + (double)(v.y * scale);
        if (accumulate) {
            C h = hs[ig];
            re += (double)h.x;
            im += (double)h.y;
        }
        hs[ig] = C{(T)re, (T)im};
    }
}

// ---------------------------------------------------------------------------
//  Split-plane kernels: grids whose n2 x n3 plane does not fit the LDS
//  (e.g. 108^3, 120^3, 128^3 in FP64).  The 2D plane transform becomes two
//  passes through the workspace (plane-major, rows in digit-reversed order
//  between the passes), each workgroup holding a chunk of rows or columns:
//    G2R  rows:  zero R rows, scatter their sphere entries (sorted by row,
//                ent_rowptr), DIT backward FFT along k, store the rows;
//         cols:  load C columns, DIT backward FFT along j, store natural;
//    R2G  cols:  load C natural columns, DIF forward FFT along j, store;
//         rows:  load R rows, DIF forward FFT along k, gather
//                hpsi = [hpsi +] kin psi + V psi / N.
//  Items are flat over (segment, plane, band, chunk), chunk fastest.
// ---------------------------------------------------------------------------
struct SplitShape {
    int rows;      // rows per row-pass workgroup
    int nrc;       // row chunks per plane
    int cols;      // columns per column-pass workgroup
    int ncc;       // column chunks per plane
};

template <typename T, bool G2R>
__global__ void __launch_bounds__(256)
k_split_rows(const LocalBatch<T> bt, SplitShape sh, int nitems, T scale, int accumulate, FftAxis a2, FftAxis a3,
             const cx_t<T>* __restrict__ tw3) {
    using C = cx_t<T>;
    const int item = blockIdx.x;
    if (item >= nitems) return;
    const LocalSeg<T>& sg = bt.s[seg_of(bt, item)];
    const int local = item - sg.item0;
    const int rc = local % sh.nrc;
    const int p = (local / sh.nrc) % sg.nplanes;
    const int b = local / sh.nrc / sg.nplanes;
    const int n2 = a2.n, n3 = a3.n;
    const int r0 = rc * sh.rows, nr = min(sh.rows, n2 - r0);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* buf = reinterpret_cast<C*>(smem_raw);          // [nr][n3]
    C* t3 = buf + (size_t)sh.rows * n3;
    const int wp = ws_pitch<T>(n3);
    C* ws = sg.work + ((int64_t)b * sg.nplanes + p) * ws_plane<T>(n2, n3) + (int64_t)r0 * wp;
    for (int i = threadIdx.x; i < n3; i += blockDim.x) t3[i] = tw3[i];
    const int* rp = sg.ent_rowptr + (size_t)p * (n2 + 1);
    const int e0 = rp[r0], e1 = rp[r0 + nr];
    const int base = r0 * n3;
    if (G2R) {
        for (int i = threadIdx.x; i < nr * n3; i += blockDim.x) buf[i] = C{0, 0};
        __syncthreads();
        const C* ps = sg.psi + (int64_t)b * sg.ld;
        for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) buf[sg.ent_pos[e] - base] = ps[sg.ent_ig[e]];
        __syncthreads();
        fft_lds<+1, true, T>(buf, a3, 1, n3, nr, t3);
        for (int i = threadIdx.x; i < nr * n3; i += blockDim.x) {
            const int r = i / n3;
            ws[(int64_t)r * wp + (i - r * n3)] = buf[i];
        }
    } else {
        for (int i = threadIdx.x; i < nr * n3; i += blockDim.x) {
            const int r = i / n3;
            buf[i] = ws[(int64_t)r * wp + (i - r * n3)];
        }
        __syncthreads();
        fft_lds<-1, false, T>(buf, a3, 1, n3, nr, t3);
        const C* ps = sg.psi + (int64_t)b * sg.ld;
        C* hs = sg.hpsi + (int64_t)b * sg.ldo;
        for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
            const int ig = sg.ent_ig[e];
            const C v = buf[sg.ent_pos[e] - base];
            const C sv = ps[ig];
            const double kn = sg.kin[ig];
            // kinetic term in FP64 (callable.cpp:119-126), then the gathered
            // local-potential term (callable.cpp:171-182)
            double re = kn * (double)sv.x + (double)(v.x * scale);
            double im = kn * (double)sv.y + (double)(v.y * scale);
            if (accumulate) {
                const C h = hs[ig];
                re += (double)h.x;
                im += (double)h.y;
            }
            hs[ig] = C{(T)re, (T)im};
        }
    }
}

// Column pass: DIT backward (S = +1, digit-reversed rows in, natural out)
// or DIF forward (S = -1, natural in, digit-reversed rows out), in place.
template <typename T, int S>
__global__ void __launch_bounds__(256)
k_split_cols(const LocalBatch<T> bt, SplitShape sh, int nitems, FftAxis a2, int n3, const cx_t<T>* __restrict__ tw2) {
    using C = cx_t<T>;
    const int item = blockIdx.x;
    if (item >= nitems) return;
    const LocalSeg<T>& sg = bt.s[seg_of(bt, item)];
    const int local = item - sg.item0;
    const int cc = local % sh.ncc;
    const int p = (local / sh.ncc) % sg.nplanes;
    const int b = local / sh.ncc / sg.nplanes;
    const int n2 = a2.n;
    const int c0 = cc * sh.cols, nc = min(sh.cols, n3 - c0);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* buf = reinterpret_cast<C*>(smem_raw);          // [n2][cols], columns adjacent
    C* t2 = buf + (size_t)n2 * sh.cols;
    const int wp = ws_pitch<T>(n3);
    C* ws = sg.work + ((int64_t)b * sg.nplanes + p) * ws_plane<T>(n2, n3) + c0;
    for (int i = threadIdx.x; i < n2; i += blockDim.x) t2[i] = tw2[i];
    for (int i = threadIdx.x; i < n2 * nc; i += blockDim.x) {
        const int j = i / nc, z = i - j * nc;
        buf[j * sh.cols + z] = ws[(int64_t)j * wp + z];
    }
    __syncthreads();
    fft_lds<S, (S > 0), T>(buf, a2, sh.cols, 1, nc, t2);
    for (int i = threadIdx.x; i < n2 * nc; i += blockDim.x) {
        const int j = i / nc, z = i - j * nc;
        ws[(int64_t)j * wp + z] = buf[j * sh.cols + z];
    }
}

// Chunk sizes of the split-plane kernels: about 64 KB of LDS per workgroup
// (two per CU), at least one row / column.
template <typename T>
static SplitShape split_shape(int n2, int n3) {
    constexpr size_t kChunkLds = 64 * 1024;
    const size_t S = sizeof(cx_t<T>);
    SplitShape sh{};
    sh.rows = (int)std::max<size_t>(1, std::min<size_t>((size_t)n2, kChunkLds / (S * n3) - 1));
    sh.nrc = (n2 + sh.rows - 1) / sh.rows;
    sh.cols = (int)std::max<size_t>(1, std::min<size_t>((size_t)n3, kChunkLds / (S * n2) - 1));
    sh.ncc = (n3 + sh.cols - 1) / sh.cols;
    return sh;
}

// ---------------------------------------------------------------------------
//  Specialised kernels for square n2 = n3 = N planes and N1-point x pencils
//  with compile-time two-stage plans (fft_ct.hpp).  Same data flow as above.
// ---------------------------------------------------------------------------
// Persistent plane kernels: each workgroup walks the flat (segment, plane,
// band) items with stride gridDim.x.  One FP64 plane fills 129.6 KB of the
// 160 KB LDS, so a CU holds one workgroup and nothing else hides the memory
// latency of an item's dependent loads (CSR offsets -> entry indices ->
// psi).  The kernels load inputs ahead, a level at a time, behind the FFT
// stages (per precision, PlaneTune):
//   SPF  G2R: the next item's row list and its first SPF scatter entries per
//        thread (G2R_EARLY below: when they load);
//   GPF  R2G: this item's first GPF gather entries per thread (indices
//        during the first column stage, psi / kin / hpsi after it);
//   YPF  R2G: column-stage inputs u < YPF of the next plane, issued before
//        this plane's row pass (in flight through the row pass and the
//        gather; the rest at the top of the item); the next row list is
//        fetched after the gather;
//   G2RW / R2GW  waves per SIMD the register allocation must allow;
//   BLOCK_CAP  largest workgroup: FP32 planes (65 KB of LDS) run two
//        workgroups of 8 waves per CU, each thread taking up to two
//        butterflies per stage, where 15-wave workgroups would need <= 64
//        VGPRs for a second one.
// Entries past SPF / GPF per thread (the densest planes) load directly.
// Measured on the 90^3 grid (256 bands): FP64 R2G 420 -> 388 us with GPF 1,
// 402 -> 385 us with YPF 4 (and 605 -> 534 us on a 48^3 grid the sphere
// fills; YPF 6-8 fit only without GPF and were slower); FP32 G2R 242 ->
// 220 us with 512-thread workgroups, FP32 R2G 236 -> 228 us at two
// workgroups per CU (either shape); YPF > 0 did not pay in FP32 (YPF 4
// with the early inputs waited for before the gather: 217-218 vs 219-222
// us; YPF 9 leaves one workgroup per CU: 261 us); SPF changes G2R by 1-3%
// with the FP64 order (the FP32 one: below).
template <typename T>
struct PlaneTune;
template <>
struct PlaneTune<double> {
    static constexpr int G2RW = 1, R2GW = 1, SPF = 1, GPF = 1, YPF = 4, BLOCK_CAP = 1024;
    static constexpr bool G2R_EARLY = false;
};
template <>
struct PlaneTune<float> {
    static constexpr int G2RW = 1, R2GW = 1, SPF = 2, GPF = 2, YPF = 0, BLOCK_CAP = 512;
    static constexpr bool G2R_EARLY = true;
};

// Workgroup size of the square-plane kernels.  The 96-point planes (12 x 8)
// have 1152 butterflies per stage: at the 1024-thread (FP64) / 512-thread
// (FP32) default two or three column butterflies per thread spill (FP64
// 128 VGPRs: 68 / 216 bytes per lane in G2R / R2G) or leave one workgroup
// per CU (FP32 166 VGPRs), so they take 576 / 384 threads (exactly two /
// three butterflies each; 170 VGPRs allowed, two FP32 workgroups per CU).
template <typename T, int N>
constexpr int plane_block_of() {
    if constexpr (N == 96) return sizeof(T) == 8 ? 576 : 384;
    return cmin(plane_block<N>(), PlaneTune<T>::BLOCK_CAP);
}

// Waits until every vector memory access this wave issued has completed
// (s_waitcnt vmcnt(0), other counters unconstrained).  A real waitcnt
// instruction: the compiler's own wait insertion counts it.
__device__ __forceinline__ void vmem_wait_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Flat item -> (segment, plane, band).
template <typename T>
struct PlaneItem {
    const LocalSeg<T>* sg;
    int p, b;
};
template <typename T>
__device__ __forceinline__ PlaneItem<T> plane_item(const LocalBatch<T>& bt, int item) {
    const LocalSeg<T>& sg = bt.s[seg_of(bt, item)];
    const int local = item - sg.item0;
    const int p = local % sg.nplanes;
    return PlaneItem<T>{&sg, p, local / sg.nplanes};
}

template <typename T, int N, int BLOCK>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(PlaneTune<T>::G2RW)))
k_plane_g2r_ct(const LocalBatch<T> bt, int nitems, const cx_t<T>* __restrict__ tw_g) {
    using C = cx_t<T>;
    constexpr int N2 = N * N;
    constexpr int WP = ws_pitch<T>(N);
    constexpr int64_t PS = ws_plane<T>(N, N);           // workspace plane stride
    constexpr int NZ = (N2 + BLOCK - 1) / BLOCK;
    constexpr int SPF = PlaneTune<T>::SPF;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* plane = reinterpret_cast<C*>(smem_raw);
    C* tw = plane + N2;
    int* srows = reinterpret_cast<int*>(tw + N);        // rows of the plane holding data
    if (threadIdx.x < N) tw[threadIdx.x] = tw_g[threadIdx.x];
    // the next item's row count, this thread's row-list entry, scatter
    // positions (-1: none) and values
    int nrows = 0, srow = 0, spos[SPF > 0 ? SPF : 1];
    C sval[SPF > 0 ? SPF : 1];
    auto fetch_rows = [&](const PlaneItem<T>& q) {
        const int r0 = q.sg->row_ptr[q.p];
        nrows = q.sg->row_ptr[q.p + 1] - r0;
        if (threadIdx.x < nrows) srow = q.sg->rows[r0 + threadIdx.x];
    };
    int sig[SPF > 0 ? SPF : 1];
    auto fetch_index = [&](const PlaneItem<T>& q) {
        const int e0 = q.sg->plane_ptr[q.p], e1 = q.sg->plane_ptr[q.p + 1];
#pragma unroll
        for (int f = 0; f < SPF; ++f) {
            const int e = e0 + threadIdx.x + f * BLOCK;
            spos[f] = -1;
            if (e < e1) {
                spos[f] = q.sg->ent_pos[e];
                sig[f] = q.sg->ent_ig[e];
            }
        }
    };
    auto fetch_values = [&](const PlaneItem<T>& q) {
        const C* ps = q.sg->psi + (int64_t)q.b * q.sg->ld;
#pragma unroll
        for (int f = 0; f < SPF; ++f)
            if (spos[f] >= 0) sval[f] = ps[sig[f]];
    };
    // EARLY (FP32): the next item's row list, scatter indices and psi values
    // load during this item's row passes and are waited for before its
    // workspace stores, and the row list goes to LDS then.  Vector loads
    // and stores share one counter, so a wait for a load issued after the
    // stores (the row-list and scatter uses at the top of the next item
    // compiled to one) waits for every store of the wave (the scatter
    // entries past SPF per thread still load after the stores: SPF = 3,
    // which covers every plane of the 90^3 grid, was slower, 195-198 vs
    // 184-186 us).  FP64 keeps the later loads: its 130 KB of stores
    // per item take longer to drain than an item's FFT stages, so any
    // wait inside the next item stalls on them, and the stalls come out
    // fewest with the loads at the end of the item.  Measured on the 90^3
    // grid (256 bands): FP32 G2R 213 -> 185 us; the same order in FP64
    // 328 -> 418 us.
    constexpr bool EARLY = PlaneTune<T>::G2R_EARLY;
    int item = blockIdx.x;
    if (item < nitems) {
        const PlaneItem<T> q = plane_item(bt, item);
        fetch_rows(q);
        fetch_index(q);
        fetch_values(q);
        if constexpr (EARLY) {
            vmem_wait_all();
            if (threadIdx.x < nrows) srows[threadIdx.x] = srow;
        }
    }
    for (; item < nitems; item += gridDim.x) {
        const PlaneItem<T> cur = plane_item(bt, item);
        const int nxt = item + gridDim.x;
        const PlaneItem<T> q = plane_item(bt, nxt < nitems ? nxt : item);
        const int nr = nrows;
        if constexpr (!EARLY)
            if (threadIdx.x < nr) srows[threadIdx.x] = srow;
#pragma unroll
        for (int it = 0; it < NZ; ++it) {
            const int i = threadIdx.x + it * BLOCK;
            if (NZ * BLOCK == N2 || i < N2) plane[i] = C{0, 0};
        }
        __syncthreads();
#pragma unroll
        for (int f = 0; f < SPF; ++f)
            if (spos[f] >= 0) plane[spos[f]] = sval[f];
        {
            const LocalSeg<T>& sg = *cur.sg;
            const C* ps = sg.psi + (int64_t)cur.b * sg.ld;
            const int e1 = sg.plane_ptr[cur.p + 1];
            for (int e = sg.plane_ptr[cur.p] + threadIdx.x + SPF * BLOCK; e < e1; e += BLOCK)
                plane[sg.ent_pos[e]] = ps[sg.ent_ig[e]];
        }
        __syncthreads();
        if (nxt < nitems) {
            fetch_rows(q);
            if constexpr (EARLY) fetch_index(q);
        }
        // along k, rows with data
        ct_stage_rows<N, Plan2<N>::P1, 1, +1, true, BLOCK, T>(plane, tw, srows, nr);
        __syncthreads();
        if (nxt < nitems) {
            if constexpr (EARLY)
                fetch_values(q);
            else
                fetch_index(q);
        }
        ct_stage_rows<N, Plan2<N>::P2, Plan2<N>::P1, +1, true, BLOCK, T>(plane, tw, srows, nr);
        __syncthreads();
        // along j (columns): first stage in LDS, the last one stores the rows
        // (pad columns of FP32 rows of 128-byte pitch carry junk)
        ct_stage<N, Plan2<N>::P1, 1, +1, true, N, 1, N, BLOCK, T>(plane, tw);
        __syncthreads();
        if (nxt < nitems) {
            if constexpr (EARLY) {
                vmem_wait_all();
                if (threadIdx.x < nrows) srows[threadIdx.x] = srow;   // this item's rows are no longer read
            } else {
                fetch_values(q);
            }
        }
        ct_cols_last_to_global<N, WP, +1, BLOCK, T>(
            plane, tw, cur.sg->work + (int64_t)cur.b * cur.sg->nplanes * PS + cur.p * cur.sg->wsp, cur.sg->wsr);
        __syncthreads();
    }
}

template <typename T, int N, int BLOCK>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(PlaneTune<T>::R2GW)))
k_plane_r2g_ct(const LocalBatch<T> bt, int nitems, T scale, int accumulate,
               const cx_t<T>* __restrict__ tw_g) {
    using C = cx_t<T>;
    constexpr int N2 = N * N;
    constexpr int WP = ws_pitch<T>(N);
    constexpr int64_t PS = ws_plane<T>(N, N);
    constexpr int GPF = PlaneTune<T>::GPF;
    using Cs = ColStage<N, N, BLOCK>;
    // column-stage inputs u < YP of the next plane load during the gather,
    // the rest at the top of the item
    constexpr int YP = PlaneTune<T>::YPF < Cs::P2 ? PlaneTune<T>::YPF : Cs::P2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* plane = reinterpret_cast<C*>(smem_raw);
    C* tw = plane + N2;
    int* srows = reinterpret_cast<int*>(tw + N);        // rows of the plane that are gathered
    if (threadIdx.x < N) tw[threadIdx.x] = tw_g[threadIdx.x];
    __syncthreads();                                    // tw is read before the first item's barrier
    // the first column stage (along j) runs on registers loaded from the
    // workspace plane, with the row list, one item ahead
    C y[Cs::ITER][Cs::P2];
    int nrows = 0, srow = 0;
    auto plane_of = [&](const PlaneItem<T>& q) {
        return static_cast<const C*>(q.sg->work) + (int64_t)q.b * q.sg->nplanes * PS + q.p * q.sg->wsp;
    };
    auto fetch_rows = [&](const PlaneItem<T>& q) {
        const int r0 = q.sg->row_ptr[q.p];
        nrows = q.sg->row_ptr[q.p + 1] - r0;
        if (threadIdx.x < nrows) srow = q.sg->rows[r0 + threadIdx.x];
    };
    auto fetch_y = [&](const PlaneItem<T>& q) {
        ct_cols_first_load<N, WP, BLOCK, T, 0, YP>(plane_of(q), y, q.sg->wsr);
    };
    int item = blockIdx.x;
    if (item < nitems) {
        fetch_rows(plane_item(bt, item));
        fetch_y(plane_item(bt, item));
    }
    for (; item < nitems; item += gridDim.x) {
        const PlaneItem<T> cur = plane_item(bt, item);
        const LocalSeg<T>& sg = *cur.sg;
        if constexpr (YP < Cs::P2) ct_cols_first_load<N, WP, BLOCK, T, YP, Cs::P2>(plane_of(cur), y, sg.wsr);
        ct_cols_first_to_lds<N, -1, BLOCK, T>(plane, tw, y);
        const int nr = nrows;
        if (threadIdx.x < nr) srows[threadIdx.x] = srow;
        __syncthreads();
        const int e0 = sg.plane_ptr[cur.p], e1 = sg.plane_ptr[cur.p + 1];
        const C* ps = sg.psi + (int64_t)cur.b * sg.ld;
        C* hs = sg.hpsi + (int64_t)cur.b * sg.ldo;
        const int* ent_ig = sg.ent_ig;
        const int* ent_pos = sg.ent_pos;
        const double* kin = sg.kin;
        // this item's first GPF gather entries: indices now, values one stage later
        constexpr int GA = GPF > 0 ? GPF : 1;
        int gi[GA], gp[GA];
        C gs[GA], gh[GA];
        double gk[GA];
#pragma unroll
        for (int f = 0; f < GPF; ++f) {
            const int e = e0 + threadIdx.x + f * BLOCK;
            gi[f] = -1;
            if (e < e1) {
                gi[f] = ent_ig[e];
                gp[f] = ent_pos[e];
            }
        }
        ct_stage<N, Cs::P1, 1, -1, false, N, 1, N, BLOCK, T>(plane, tw);   // along j, last stage
        __syncthreads();
#pragma unroll
        for (int f = 0; f < GPF; ++f) {
            if (gi[f] >= 0) {
                gs[f] = ps[gi[f]];
                gk[f] = kin[gi[f]];
                gh[f] = accumulate ? hs[gi[f]] : C{0, 0};
            }
        }
        const int nxt = item + gridDim.x;
        if (nxt < nitems) fetch_y(plane_item(bt, nxt));
        ct_fft_rows<N, -1, false, BLOCK, T>(plane, tw, srows, nr);        // along k, gathered rows
#pragma unroll
        for (int f = 0; f < GPF; ++f) {
            if (gi[f] < 0) continue;
            const C v = plane[gp[f]];
            const double re = gk[f] * (double)gs[f].x + (double)(v.x * scale) + (double)gh[f].x;
            const double im = gk[f] * (double)gs[f].y + (double)(v.y * scale) + (double)gh[f].y;
            hs[gi[f]] = C{(T)re, (T)im};
        }
#pragma unroll 1
        for (int e = e0 + threadIdx.x + GPF * BLOCK; e < e1; e += BLOCK) {
            const int ig = ent_ig[e];
            const C v = plane[ent_pos[e]];
            const C sv = ps[ig];
            double re = kin[ig] * (double)sv.x + (double)(v.x * scale);
            double im = kin[ig] * (double)sv.y + (double)(v.y * scale);
            if (accumulate) {
                const C h = hs[ig];
                re += (double)h.x;
                im += (double)h.y;
            }
            hs[ig] = C{(T)re, (T)im};
        }
        if (nxt < nitems) fetch_rows(plane_item(bt, nxt));
        __syncthreads();
    }
}

// x pencils of one (j, z-chunk) for NBW consecutive bands of one segment,
// as register FFTs.  N1 = P1 * P2 (Plan2, P1 >= P2).  Thread (z, role):
//   G2R (DIT)  role b < P2 loads x = P2 u + b (u < P1) straight from the
//              plane workspace, DFT_P1 -> LDS transpose -> role k1 < P1
//              gathers Z[u'][k1], twiddle W^(u' k1), DFT_P2: it now holds
//              natural x = k1 + P1 q' (q' < P2);
//   V          multiply by V_loc(x, j, z) kept in registers across bands;
//   R2G (DIF)  DFT_P2 on the same values, twiddle W^(q k1) -> LDS transpose
//              -> role b gathers, DFT_P1 -> holds x = P2 q + b, stored to
//              the active planes only.
// Two LDS round trips per band through one N1 x TZ buffer, two barriers:
// each thread's LDS writes of a phase go to exactly the slots it read in the
// phase before ((u P1 + k1) for role k1 in the middle phase, (b P1 + q) for
// role b in the outer ones), so no write can overtake another thread's read
// and one buffer serves all three phases.
// The kernel is VALU-bound in FP32, so the per-band instruction count is
// kept close to the FFT itself: the x -> (plane) byte offsets and active
// flags of a thread's P1 pencil points are band-invariant and computed once
// (HOIST; a band's loads and stores are then a uniform base plus a 32-bit
// offset), and twiddle W^0 = 1 is multiplied rather than selected.  The
// loads of a band are not issued ahead of the previous band's arithmetic
// (BPF): without that register set the kernel fits 126 VGPRs in FP64 and
// 81 in FP32, i.e. 5 resident workgroups per CU instead of 4 (FP64) and 8
// instead of 6 (FP32), which hide the latency better.  Measured on 90^3
// (256 bands): FP32 429 us -> 358 us (HOIST) -> 348 us (no BPF); FP64
// 665 us -> 594 us (HOIST, no BPF; with BPF, hoisting costs a wave per
// SIMD and 746 us); V_loc in LDS instead of registers (VR) was slower.
// Running two bands per loop trip on two register sets needs about 60 more
// VGPRs and was slower in both precisions.
// Grid: (n2 * z-chunks, nseg * band groups).
template <typename T, int N1>
struct PencilShape;

template <typename T, int N1, int TZ, int NBW>
__global__ void __launch_bounds__(TZ * Plan2<N1>::P1)
__attribute__((amdgpu_waves_per_eu(PencilShape<T, N1>::W)))
k_xpencil_vloc_ct(const LocalBatch<T> bt, const T* __restrict__ vloc, int n2, int n3,
                  const cx_t<T>* __restrict__ tw_g) {
    using C = cx_t<T>;
    constexpr int P1 = Plan2<N1>::P1, P2 = Plan2<N1>::P2;
    static_assert(P1 >= P2, "roles = P1");
    __shared__ C ex[N1 * TZ];
    __shared__ C tw[N1];
    // logical workgroup (z-chunk fastest, then row j, band group, segment),
    // in XCD order: the z-chunks of one row share cache lines at their
    // edges, and on one XCD they share its L2 (measured on the 64-atom step:
    // FP64 fetch 1.98 -> 1.48 GB per launch against 1.36 GB read; time
    // unchanged, the pass runs at the device's copy rate)
    const int G = (int)(gridDim.x * gridDim.y);
    const int L = xcd_order((int)(blockIdx.x + blockIdx.y * gridDim.x), G);
    const int bx = L % (int)gridDim.x, by = L / (int)gridDim.x;
    const int nbg = (bt.nb + NBW - 1) / NBW;
    const int sidx = by / nbg;
    const int bgrp = by - sidx * nbg;
    const LocalSeg<T>& sg = bt.s[sidx];
    const int nzc = (n3 + TZ - 1) / TZ;
    const int j = bx / nzc;
    const int z0 = (bx - j * nzc) * TZ;
    const int z = threadIdx.x % TZ;
    const int role = threadIdx.x / TZ;
    const bool zv = z0 + z < n3;
    const int64_t n23 = (int64_t)n2 * n3;
    // With gcd(P1, P2) = 1 (90 = 10 x 9, 40, 45, 20, 15) the prime-factor
    // mapping replaces the Cooley-Tukey one and the twiddles between the
    // stages disappear: role b < P2 loads / stores x = (P2 u + P1 b) mod N1
    // instead of P2 u + b, and role k1 holds x = (P2 a k1 + P1 c q) mod N1
    // (a = P2^-1 mod P1, c = P1^-1 mod P2) instead of k1 + P1 q between the
    // stages.  Measured on 90^3, 256 bands: FP32 pass 333 -> 305 us, FP64
    // 578 -> 574 us; the mixed R2G pass after it took 234 -> 249 us in the
    // same runs; step 1.385 -> 1.367 ms mixed, 2.55 -> 2.53 ms FP64.
    constexpr bool PFA = gcd_c(P1, P2) == 1;
    constexpr int PA = inv_mod(P2 % P1, P1), PC = inv_mod(P1 % P2, P2);
    auto xin = [&](int u) { return PFA ? (P2 * u + P1 * role) % N1 : P2 * u + role; };
    auto xmid = [&](int q) { return PFA ? (P2 * PA * role + P1 * PC * q) % N1 : role + P1 * q; };
    if constexpr (!PFA)
        for (int i = threadIdx.x; i < N1; i += TZ * P1) tw[i] = tw_g[i];
    // V at the natural x = xmid(q) this thread scales: in registers
    // (VR) or in LDS (x * TZ + z), which frees 2 P2 FP64 registers
    constexpr bool VR = PencilShape<T, N1>::VR;
    __shared__ T vs[VR ? 1 : N1 * TZ];
    T vr[P2];
    if constexpr (VR) {
#pragma unroll
        for (int q = 0; q < P2; ++q) {
            const int x = xmid(q);
            vr[q] = zv ? vloc[(int64_t)x * n23 + (int64_t)j * n3 + z0 + z] : T(0);
        }
    } else {
        for (int i = threadIdx.x; i < N1 * TZ; i += TZ * P1) {
            const int x = i / TZ, zz = i - x * TZ;
            vs[i] = z0 + zz < n3 ? vloc[(int64_t)x * n23 + (int64_t)j * n3 + z0 + zz] : T(0);
        }
    }
    // this thread's pencil points x = P2 u + role: byte offset inside a
    // band's workspace and whether the plane is stored (active) at all
    const int64_t wps = ws_plane<T>(n2, n3);
    const int64_t band_bytes = (int64_t)sg.nplanes * wps * (int64_t)sizeof(C);
    const unsigned colb = (unsigned)(((int64_t)j * sg.wsr + z0 + z) * (int64_t)sizeof(C));
    // (inactive: offset kNone).  HOIST keeps them in registers across the
    // bands; otherwise each band recomputes them from the plane map in LDS.
    constexpr bool HOIST = PencilShape<T, N1>::HOIST;
    constexpr unsigned kNone = ~0u;
    __shared__ int srcpl[HOIST ? 1 : N1];
    const unsigned plane_bytes = (unsigned)(sg.wsp * (int64_t)sizeof(C));
    unsigned boff[P1];
    bool actr[P1];
    if constexpr (HOIST) {
#pragma unroll
        for (int u = 0; u < P1; ++u) {
            const int pl = (role < P2 && zv) ? sg.x_to_plane[xin(u)] : -1;
            actr[u] = pl >= 0;
            boff[u] = pl >= 0 ? (unsigned)pl * plane_bytes + colb : kNone;
        }
    } else {
        for (int i = threadIdx.x; i < N1; i += TZ * P1) srcpl[i] = sg.x_to_plane[i];
    }
    auto off = [&](int u) {
        if constexpr (HOIST) return boff[u];
        const int pl = (role < P2 && zv) ? srcpl[xin(u)] : -1;
        return pl >= 0 ? (unsigned)pl * plane_bytes + colb : kNone;
    };
    auto act = [&](int u, unsigned o) { return HOIST ? actr[u] : o != kNone; };
    __syncthreads();
    const char* work = reinterpret_cast<const char*>(sg.work);
    auto load = [&](int b, C (&y)[P1]) {
        const char* wb = work + (int64_t)b * band_bytes;
#pragma unroll
        for (int u = 0; u < P1; ++u) {
            const unsigned o = off(u);
            y[u] = act(u, o) ? *reinterpret_cast<const C*>(wb + o) : C{0, 0};
        }
    };
    auto band = [&](int b, C (&y)[P1]) {
        // ---- G2R stage 1: DFT_P1 over u, positions b*P1 + q
        if (role < P2) {
            dft<P1, +1, T>(y);
#pragma unroll
            for (int q = 0; q < P1; ++q) ex[(role * P1 + q) * TZ + z] = y[q];
        }
        __syncthreads();
        {
            // ---- G2R stage 2: twiddle W^(u' k1) (backward), DFT_P2
            C w[P2];
#pragma unroll
            for (int u = 0; u < P2; ++u) {
                const C a = ex[(u * P1 + role) * TZ + z];
                w[u] = (PFA || u == 0) ? a : cmul(a, twid<+1>(tw, u * role));
            }
            dft<P2, +1, T>(w);
            // ---- V_loc(r) at natural x = xmid(q)
#pragma unroll
            for (int q = 0; q < P2; ++q) {
                const T vq = VR ? vr[q] : vs[xmid(q) * TZ + z];
                w[q] = C{w[q].x * vq, w[q].y * vq};
            }
            // ---- R2G stage 1 (DIF): DFT_P2, twiddle W^(q k1) (forward)
            dft<P2, -1, T>(w);
#pragma unroll
            for (int q = 0; q < P2; ++q)
                ex[(q * P1 + role) * TZ + z] = (PFA || q == 0) ? w[q] : cmul(w[q], twid<-1>(tw, q * role));
        }
        __syncthreads();
        if (role < P2) {
            // ---- R2G stage 2: DFT_P1, natural x = P2 q + b
#pragma unroll
            for (int u = 0; u < P1; ++u) y[u] = ex[(role * P1 + u) * TZ + z];
            dft<P1, -1, T>(y);
            char* wb = const_cast<char*>(work) + (int64_t)b * band_bytes;
#pragma unroll
            for (int q = 0; q < P1; ++q) {
                const unsigned o = off(q);
                if (act(q, o)) *reinterpret_cast<C*>(wb + o) = y[q];
            }
        }
    };
    const int bfirst = bgrp * NBW;
    const int bend = min(bt.nb, bfirst + NBW);
    // BPF: the next band's loads fly during this band's arithmetic (P1 more
    // complex registers); otherwise more resident workgroups hide them
    C v[P1];
    if constexpr (PencilShape<T, N1>::BPF) {
        C y[P1];
        if (bfirst < bend) load(bfirst, y);
        for (int b = bfirst; b < bend; ++b) {
#pragma unroll
            for (int u = 0; u < P1; ++u) v[u] = y[u];
            if (b + 1 < bend) load(b + 1, y);
            band(b, v);
        }
    } else {
        for (int b = bfirst; b < bend; ++b) {
            load(b, v);
            band(b, v);
        }
    }
}

// default z-chunk width of the pencil kernel, one of kPencilZChoices
// (measured: FP32 prefers chunks of one 128-byte line, FP64 on the
// 90-point grid five chunks of 18); effm_autotune times both widths
template <typename T, int N> struct PencilShape {
    static constexpr int TZ = 16, W = 1;
    static constexpr bool VR = true, HOIST = true, BPF = false;
};
template <> struct PencilShape<double, 90> {
    static constexpr int TZ = 18, W = 1;
    static constexpr bool VR = true, HOIST = true, BPF = false;
};


static int Plan2N(int n, int which) {
    switch (n) {
#define EFFM_PLAN_CASE(N) \
    case N: return which == 1 ? Plan2<N>::P1 : Plan2<N>::P2;
        EFFM_CT_SIZES(EFFM_PLAN_CASE)
#undef EFFM_PLAN_CASE
        default: return -1;
    }
}

static bool plan_is(const FftAxis& a, int p1, int p2) {
    return a.ns == 2 && a.p[0] == p1 && a.p[1] == p2;
}

// The square-plane kernels with compile-time plans serve this cell, with
// coprime radices: their scatter / gather positions are then the
// prime-factor order of Pfa<N> (p1, p2 = the radices).
bool plane_ct_pfa(const Ctx& c, int* p1, int* p2) {
    const int a = Plan2N(c.n2, 1), b = Plan2N(c.n2, 2);
    if (c.n2 != c.n3 || a <= 0 || !plan_is(c.ax2, a, b) || !plan_is(c.ax3, a, b)) return false;
    int x = a, y = b;
    while (y) { const int t = x % y; x = y; y = t; }
    if (x != 1) return false;
    *p1 = a;
    *p2 = b;
    return true;
}

// Persistent grid: as many workgroups as the occupancy calculator says are
// resident at once (LDS, VGPRs, wave slots), capped by the items.  The
// dynamic-LDS attribute and the occupancy are set up once per kernel and
// context.
static hipError_t persistent_grid(Ctx& c, const void* fn, size_t lds, int block, int nitems, int* grid) {
    auto it = c.occupancy.find(fn);
    if (it == c.occupancy.end()) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, lds) != hipSuccess || per_cu < 1)
            per_cu = 1;
        it = c.occupancy.emplace(fn, per_cu).first;
    }
    *grid = std::max(1, std::min(nitems, c.ncu * it->second));
    return hipSuccess;
}

template <typename T, int N>
static hipError_t launch_planes_ct(Ctx& c, const LocalBatch<T>& bt, int nitems, bool accumulate,
                                   bool g2r, const cx_t<T>* tw) {
    constexpr int BLOCK = plane_block_of<T, N>();
    const size_t lds = sizeof(cx_t<T>) * ((size_t)N * N + N) + sizeof(int) * N;
    hipError_t e;
    int grid = 0;
    if (g2r) {
        e = persistent_grid(c, reinterpret_cast<const void*>(&k_plane_g2r_ct<T, N, BLOCK>), lds, BLOCK,
                            nitems, &grid);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_plane_g2r_ct<T, N, BLOCK>), dim3(grid), dim3(BLOCK), lds, c.stream, bt,
                           nitems, tw);
    } else {
        e = persistent_grid(c, reinterpret_cast<const void*>(&k_plane_r2g_ct<T, N, BLOCK>), lds, BLOCK,
                            nitems, &grid);
        if (e != hipSuccess) return e;
        T scale = (T)(1.0 / ((double)c.n1 * c.n2 * c.n3));
        hipLaunchKernelGGL((k_plane_r2g_ct<T, N, BLOCK>), dim3(grid), dim3(BLOCK), lds, c.stream, bt,
                           nitems, scale, accumulate ? 1 : 0, tw);
    }
    return hipGetLastError();
}

template <typename T, int N, int TZ, int NBW>
static void launch_pencil_ct_nbw(Ctx& c, const LocalBatch<T>& bt, const T* vl, const cx_t<T>* tw1) {
    constexpr int BLOCK = TZ * cmax(Plan2<N>::P1, Plan2<N>::P2);
    const int nzc = (c.n3 + TZ - 1) / TZ;
    const int nbg = (bt.nb + NBW - 1) / NBW;
    dim3 gx(c.n2 * nzc, nbg * bt.nseg);
    hipLaunchKernelGGL((k_xpencil_vloc_ct<T, N, TZ, NBW>), gx, dim3(BLOCK), 0, c.stream, bt, vl, c.n2, c.n3, tw1);
}

template <typename T, int N, int TZ>
static void launch_pencil_ct_z(Ctx& c, const LocalBatch<T>& bt, const T* vl, const cx_t<T>* tw1) {
    switch (c.tuning<T>().pencil_bands) {
        case 4: launch_pencil_ct_nbw<T, N, TZ, 4>(c, bt, vl, tw1); break;
        case 16: launch_pencil_ct_nbw<T, N, TZ, 16>(c, bt, vl, tw1); break;
        default: launch_pencil_ct_nbw<T, N, TZ, 8>(c, bt, vl, tw1); break;
    }
}

// z points and bands per workgroup: the tuned values (kPencilZChoices,
// kPencilBandChoices)
template <typename T, int N>
static void launch_pencil_ct(Ctx& c, const LocalBatch<T>& bt, const T* vl, const cx_t<T>* tw1) {
    static_assert(PencilShape<T, N>::TZ == 16 || PencilShape<T, N>::TZ == 18, "a candidate width");
    const int tz = c.tuning<T>().pencil_z ? c.tuning<T>().pencil_z : PencilShape<T, N>::TZ;
    if (tz == 18) launch_pencil_ct_z<T, N, 18>(c, bt, vl, tw1);
    else launch_pencil_ct_z<T, N, 16>(c, bt, vl, tw1);
}

template <typename T>
int pencil_z_in_effect(const Ctx& c) {
    if (c.tuning<T>().pencil_z) return c.tuning<T>().pencil_z;
    switch (c.n1) {
#define EFFM_PENCIL_Z_CASE(N) \
    case N: return PencilShape<T, N>::TZ;
        EFFM_CT_SIZES(EFFM_PENCIL_Z_CASE)
#undef EFFM_PENCIL_Z_CASE
        default: return 16;
    }
}
template int pencil_z_in_effect<double>(const Ctx&);
template int pencil_z_in_effect<float>(const Ctx&);

template <typename T>
static hipError_t planes_ct(Ctx& c, int n, const LocalBatch<T>& bt, int nitems, bool accumulate, bool g2r,
                            const cx_t<T>* tw) {
    switch (n) {
#define EFFM_PLANES_CASE(N) \
    case N: return launch_planes_ct<T, N>(c, bt, nitems, accumulate, g2r, tw);
        EFFM_CT_SIZES(EFFM_PLANES_CASE)
#undef EFFM_PLANES_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
static void pencil_ct(Ctx& c, int n, const LocalBatch<T>& bt, const T* vl, const cx_t<T>* tw1) {
    switch (n) {
#define EFFM_PENCIL_CASE(N) \
    case N: launch_pencil_ct<T, N>(c, bt, vl, tw1); break;
        EFFM_CT_SIZES(EFFM_PENCIL_CASE)
#undef EFFM_PENCIL_CASE
        default: break;
    }
}

template <typename T>
static size_t plane_lds_bytes(const Ctx& c) {
    return plane_lds_bytes<T>(c.n2, c.n3);
}

template <typename T>
size_t local_work_bytes(const Ctx& c, const ApplySeg* segs, int nseg, int nb) {
    size_t tot = 0;
    for (int i = 0; i < nseg; ++i)
        tot += (size_t)segs[i].k->nplanes * ws_plane<T>(c.n2, c.n3) * nb * sizeof(cx_t<T>);
    return tot;
}

// Workspace of the compile-time kernels row-major ([band][row][plane][z])
// rather than plane-major ([band][plane][row][z]).
constexpr bool kRowMajorWork = true;

// Split-plane sequence (planes beyond the LDS): rows, columns, x pencils,
// columns, rows + gather, all on plane-major workspace.
template <typename T>
static hipError_t launch_local_split(Ctx& c, LocalBatch<T> bt, int nb, bool accumulate, const cx_t<T>* tw1,
                                     const cx_t<T>* tw2, const cx_t<T>* tw3, const T* vl) {
    using C = cx_t<T>;
    constexpr int TZ = 16;
    const SplitShape sh = split_shape<T>(c.n2, c.n3);
    LocalBatch<T> br = bt, bcl = bt;
    int nr_items = 0, nc_items = 0;
    for (int i = 0; i < bt.nseg; ++i) {
        br.s[i].item0 = nr_items;
        bcl.s[i].item0 = nc_items;
        nr_items += bt.s[i].nplanes * nb * sh.nrc;
        nc_items += bt.s[i].nplanes * nb * sh.ncc;
    }
    const size_t lds_rows = sizeof(C) * ((size_t)sh.rows * c.n3 + c.n3);
    const size_t lds_cols = sizeof(C) * ((size_t)c.n2 * sh.cols + c.n2);
    const size_t lds_pencil = sizeof(C) * ((size_t)c.n1 * TZ + c.n1);
    hipError_t e;
    const void* fns[] = {reinterpret_cast<const void*>(&k_split_rows<T, true>),
                         reinterpret_cast<const void*>(&k_split_rows<T, false>),
                         reinterpret_cast<const void*>(&k_split_cols<T, +1>),
                         reinterpret_cast<const void*>(&k_split_cols<T, -1>),
                         reinterpret_cast<const void*>(&k_xpencil_vloc<T, TZ>)};
    const size_t ldss[] = {lds_rows, lds_rows, lds_cols, lds_cols, lds_pencil};
    for (int f = 0; f < 5; ++f)
        if ((e = hipFuncSetAttribute(fns[f], hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldss[f])) != hipSuccess)
            return e;
    const T scale = (T)(1.0 / ((double)c.n1 * c.n2 * c.n3));
    timer_begin(c);
    hipLaunchKernelGGL((k_split_rows<T, true>), dim3(nr_items), dim3(256), lds_rows, c.stream, br, sh, nr_items,
                       scale, 0, c.ax2, c.ax3, tw3);
    hipLaunchKernelGGL((k_split_cols<T, +1>), dim3(nc_items), dim3(256), lds_cols, c.stream, bcl, sh, nc_items, c.ax2,
                       c.n3, tw2);
    timer_end(c, "plane_g2r");
    if ((e = hipGetLastError()) != hipSuccess) return e;
    timer_begin(c);
    const int nzc = (c.n3 + TZ - 1) / TZ;
    hipLaunchKernelGGL((k_xpencil_vloc<T, TZ>), dim3(c.n2 * nzc, nb, bt.nseg), dim3(256), lds_pencil, c.stream, bt,
                       static_cast<const int*>(c.perm1.ptr), vl, c.ax1, c.n2, c.n3, tw1);
    timer_end(c, "xpencil_vloc");
    if ((e = hipGetLastError()) != hipSuccess) return e;
    timer_begin(c);
    hipLaunchKernelGGL((k_split_cols<T, -1>), dim3(nc_items), dim3(256), lds_cols, c.stream, bcl, sh, nc_items, c.ax2,
                       c.n3, tw2);
    hipLaunchKernelGGL((k_split_rows<T, false>), dim3(nr_items), dim3(256), lds_rows, c.stream, br, sh, nr_items,
                       scale, accumulate ? 1 : 0, c.ax2, c.ax3, tw3);
    timer_end(c, "plane_r2g");
    return hipGetLastError();
}

template <typename T>
hipError_t launch_local(Ctx& c, const ApplySeg* segs, int nseg, int b0, int nb, bool accumulate,
                        void* work) {
    using C = cx_t<T>;
    constexpr int TZ = 16;
    if (nseg < 1 || nseg > kMaxSeg || nb < 1) return hipErrorInvalidValue;
    // compile-time plans for square planes / pencil lengths of EFFM_CT_SIZES
    const bool plct = c.n2 == c.n3 && Plan2N(c.n2, 1) > 0 && plan_is(c.ax2, Plan2N(c.n2, 1), Plan2N(c.n2, 2)) &&
                      plan_is(c.ax3, Plan2N(c.n3, 1), Plan2N(c.n3, 2));
    // (the specialised pencil kernel addresses one band's workspace with
    // 32-bit byte offsets)
    const bool pect = Plan2N(c.n1, 1) > 0 && plan_is(c.ax1, Plan2N(c.n1, 1), Plan2N(c.n1, 2)) &&
                      (int64_t)c.n1 * ws_plane<T>(c.n2, c.n3) * (int64_t)sizeof(C) < (int64_t{1} << 32);
    // row-major workspace when the compile-time plane and pencil kernels
    // both run (they alone read the strides)
    const bool rowmaj = kRowMajorWork && plct && pect;
    LocalBatch<T> bt{};
    bt.nseg = nseg;
    bt.nb = nb;
    int nitems = 0, maxplanes = 0;
    C* wcur = static_cast<C*>(work);
    for (int i = 0; i < nseg; ++i) {
        const KptDev& k = *segs[i].k;
        LocalSeg<T>& s = bt.s[i];
        s.psi = static_cast<const C*>(segs[i].psi) + (int64_t)b0 * segs[i].ld;
        s.hpsi = static_cast<C*>(segs[i].out[0]) + (int64_t)b0 * segs[i].ldo;
        s.ld = segs[i].ld;
        s.ldo = segs[i].ldo;
        s.plane_ptr = static_cast<const int*>(k.plane_ptr.ptr);
        s.ent_ig = static_cast<const int*>(k.ent_ig.ptr);
        s.ent_pos = static_cast<const int*>(k.ent_pos.ptr);
        s.x_to_plane = static_cast<const int*>(k.x_to_plane.ptr);
        s.plane_x = static_cast<const int*>(k.plane_x.ptr);
        s.row_ptr = static_cast<const int*>(k.row_ptr.ptr);
        s.rows = static_cast<const int*>(k.rows.ptr);
        s.kin = static_cast<const double*>(k.kin.ptr);
        s.ent_rowptr = static_cast<const int*>(k.ent_rowptr.ptr);
        s.work = wcur;
        s.nplanes = k.nplanes;
        s.wsp = rowmaj ? (int64_t)ws_pitch<T>(c.n3) : ws_plane<T>(c.n2, c.n3);
        s.wsr = rowmaj ? (int64_t)k.nplanes * ws_pitch<T>(c.n3) : (int64_t)ws_pitch<T>(c.n3);
        s.item0 = nitems;
        wcur += (int64_t)k.nplanes * ws_plane<T>(c.n2, c.n3) * nb;
        nitems += k.nplanes * nb;
        maxplanes = std::max(maxplanes, k.nplanes);
    }
    if (nitems == 0) return hipSuccess;
    const bool f64 = sizeof(T) == 8;
    const C* tw1 = static_cast<const C*>(f64 ? c.tw1_64.ptr : c.tw1_32.ptr);
    const C* tw2 = static_cast<const C*>(f64 ? c.tw2_64.ptr : c.tw2_32.ptr);
    const C* tw3 = static_cast<const C*>(f64 ? c.tw3_64.ptr : c.tw3_32.ptr);
    const T* vl = static_cast<const T*>(f64 ? c.vloc64.ptr : c.vloc32.ptr);
    const size_t lds_plane = plane_lds_bytes<T>(c);
    const size_t lds_pencil = sizeof(C) * ((size_t)c.n1 * TZ + c.n1);
    hipError_t e;

    if (!plct && lds_plane > kPlaneLdsCap) {
        for (int i = 0; i < nseg; ++i)
            if (!segs[i].k->ent_rowptr.ptr) return hipErrorInvalidValue;   // built at upload for such grids
        return launch_local_split<T>(c, bt, nb, accumulate, tw1, tw2, tw3, vl);
    }
    if (!plct || !pect) {
        // generic kernels: dynamic LDS sized by this cell's grid
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_plane_g2r<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_plane);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_plane_r2g<T>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_plane);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_xpencil_vloc<T, TZ>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_pencil);
        if (e != hipSuccess) return e;
    }
    const dim3 gp(maxplanes, nb, nseg);
    timer_begin(c);
    if (plct) e = planes_ct<T>(c, c.n2, bt, nitems, accumulate, true, tw2);
    else {
        hipLaunchKernelGGL((k_plane_g2r<T>), gp, dim3(512), lds_plane, c.stream, bt, c.ax2, c.ax3, tw2, tw3);
        e = hipGetLastError();
    }
    timer_end(c, "plane_g2r");
    if (e != hipSuccess) return e;
    timer_begin(c);
    if (pect) pencil_ct<T>(c, c.n1, bt, vl, tw1);
    else {
        const int nzc = (c.n3 + TZ - 1) / TZ;
        hipLaunchKernelGGL((k_xpencil_vloc<T, TZ>), dim3(c.n2 * nzc, nb, nseg), dim3(256), lds_pencil, c.stream,
                           bt, static_cast<const int*>(c.perm1.ptr), vl, c.ax1, c.n2, c.n3, tw1);
    }
    timer_end(c, "xpencil_vloc");
    if ((e = hipGetLastError()) != hipSuccess) return e;
    timer_begin(c);
    if (plct) e = planes_ct<T>(c, c.n2, bt, nitems, accumulate, false, tw2);
    else {
        T scale = (T)(1.0 / ((double)c.n1 * c.n2 * c.n3));
        hipLaunchKernelGGL((k_plane_r2g<T>), gp, dim3(512), lds_plane, c.stream, bt, scale, accumulate ? 1 : 0,
                           c.ax2, c.ax3, tw2, tw3);
        e = hipGetLastError();
    }
    timer_end(c, "plane_r2g");
    return e;
}

template hipError_t launch_local<double>(Ctx&, const ApplySeg*, int, int, int, bool, void*);
template hipError_t launch_local<float>(Ctx&, const ApplySeg*, int, int, int, bool, void*);
template size_t local_work_bytes<double>(const Ctx&, const ApplySeg*, int, int);
template size_t local_work_bytes<float>(const Ctx&, const ApplySeg*, int, int);

}  // namespace effm
