/*
 * effm_hpsi.h — C-ABI of libeffm_hip.so, the MI355X (gfx950) H|psi> engine.
 *
 * Plain C: pointers, sizes and status codes, no C++ or torch types.  The
 * reference has no FFI (it is a C++23 module library); each entry point below
 * replaces the reference interface cited beside it, and the C++ / Python
 * mirrors of the reference classes (effmass_amd/host/effm_hamiltonian.hpp,
 * effmass_amd/hamiltonian.py) are thin layers over these calls.
 *
 * Conventions (identical to the reference):
 *   - Hartree atomic units; K = k + G in Cartesian bohr^-1;
 *   - psi / hpsi are band-major: band b starts at psi + b*ld (complex
 *     elements), G in the caller's (GKK) order;
 *   - complex numbers are interleaved (re, im): complex128 for EFFM_PREC_FP64,
 *     complex64 for EFFM_PREC_MIXED;
 *   - without EFFM_F_OVERWRITE results are ACCUMULATED (hpsi += H psi), as
 *     Hamiltonian::Callable::operator() does
 *     (src/H_psi/hamiltonian_callable.cpp:87-92);
 *   - array arguments of effm_hpsi / effm_dhdk / effm_rayleigh are device
 *     pointers unless EFFM_F_HOST is set;
 *   - all work is ordered on the context's stream (effm_set_stream).
 */
#ifndef EFFM_HPSI_H
#define EFFM_HPSI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: map onto the reference's exception classes ---------- */
#define EFFM_OK                 0
#define EFFM_ERR_RUNTIME       -1   /* std::runtime_error  */
#define EFFM_ERR_OUT_OF_RANGE  -2   /* std::out_of_range   */
#define EFFM_ERR_DOMAIN        -3   /* std::domain_error   */
#define EFFM_ERR_INVALID       -4   /* std::invalid_argument */
#define EFFM_ERR_NOMEM         -5   /* HBM budget or hipMalloc failure */
#define EFFM_ERR_HIP           -6   /* any other HIP runtime error */

/* ---- precision ---------------------------------------------------------- */
#define EFFM_PREC_FP64   0   /* complex128 psi; FP64 FFT, GEMM, reductions */
#define EFFM_PREC_MIXED  1   /* complex64 psi; FP32 FFT + GEMM with FP32
                                accumulation; FP64 D contraction, Rayleigh */

/* ---- flags -------------------------------------------------------------- */
#define EFFM_F_OVERWRITE  0x1  /* out = H psi instead of out += H psi */
#define EFFM_F_HOST       0x2  /* psi / outputs are host pointers */
#define EFFM_F_GRAPH      0x4  /* replay a hipGraph of this apply: captured on
                                  first use, keyed by k-points, pointers,
                                  leading dimensions, nb, precision, flags
                                  and stream; device pointers only */

typedef struct effm_ctx effm_ctx;

/* One pseudopotential species after sortByL + diagonalizeNonlocal.
 * Replaces the NCPP + BetaqTables / DBetaqTables pair the Hamiltonian holds
 * per element (src/H_psi/hamiltonian.cppm:325;
 * src/pseudo/ncpp_advance.cppm:522-701, 718-895). */
typedef struct {
    int32_t lmax;              /* -1: no projectors */
    int32_t nbeta;             /* radial channels, sorted by l */
    const int32_t* beta_l;     /* [nbeta] */
    const double* lambda;      /* [nbeta] eigenvalues of the l-blocks of D (Ha) */
    int32_t nq;                /* table nodes q_i = i*dq */
    double dq;
    double qmax;
    const double* beta_tab;    /* [nbeta*nq] raw beta(q_i), no 1/sqrt(Omega) */
    const double* dbeta_tab;   /* [nbeta*nq] raw dbeta/dq(q_i), or NULL */
} effm_species;

/* The k-independent cell data: replaces Hamiltonian::finalize()'s products
 * (src/H_psi/hamiltonian.cpp:267-392): FFT grid, V_loc (VR, src/io/RHO.cppm:
 * 57-61), atoms in ascending-Z / stable order (src/io/ATOM.cppm:290-316),
 * Omega (src/io/lattice.cppm:169-190). */
typedef struct {
    int32_t n1, n2, n3;
    double lattice[9];         /* rows a1, a2, a3 (bohr) */
    const double* vloc;        /* [n1*n2*n3] Hartree, index i*n2*n3 + j*n3 + k */
    int32_t natom;
    const int32_t* species;    /* [natom] index into species[], grouped */
    const double* tau;         /* [natom*3] fractional */
    int32_t nspecies;
    const effm_species* sp;    /* [nspecies] */
} effm_cell;

/* Context: device, stream, HBM budget (bytes; 0 = no limit).
 * Replaces: construction of H_psi::Hamiltonian (src/H_psi/hamiltonian.cppm:119). */
int effm_ctx_create(effm_ctx** out, int device, uint64_t hbm_budget);
int effm_ctx_destroy(effm_ctx* ctx);
/* Order all later work after / on this HIP stream (hipStream_t as void*). */
int effm_set_stream(effm_ctx* ctx, void* stream);
const char* effm_last_error(const effm_ctx* ctx);
/* Device bytes held by the context and the budget. */
int effm_mem_info(const effm_ctx* ctx, uint64_t* used, uint64_t* budget);

/* Upload the cell.  Replaces Hamiltonian::loadVR/loadATOM/loadNCPP +
 * finalize() (src/H_psi/hamiltonian.cpp:131-153, 267-392).  Drops all
 * k-point data. */
int effm_cell_upload(effm_ctx* ctx, const effm_cell* cell);

/* Upload one k-point's G-sphere.  Replaces GKK::loadKPoint + the
 * Callable's set_ikpt (src/io/GKK.cppm:511-593,
 * src/H_psi/hamiltonian_callable.cpp:62-85).  K: [npw*3] Cartesian k+G,
 * kinetic: [npw] (validated against |K|^2/2 to 1e-10(1+E) as
 * GKK::validateKineticConsistency, src/io/GKK.cppm:652-671), gidx: [npw*3]
 * centred integer G, kfrac: fractional k.  Host pointers. */
int effm_kpt_upload(effm_ctx* ctx, int ikpt, int npw, const double* K,
                    const double* kinetic, const int32_t* gidx,
                    const double kfrac[3]);
/* Number of G vectors of a k-point (Callable::dim, hamiltonian.cppm:184). */
int effm_kpt_dim(const effm_ctx* ctx, int ikpt);
int effm_kpt_release(effm_ctx* ctx, int ikpt);

/* H|psi> for nb bands at k-point ikpt.
 * Replaces Hamiltonian::Callable::operator()(psi, hpsi) and the declared but
 * undefined batched operator()(n_bands, psi, hpsi)
 * (src/H_psi/hamiltonian.cppm:177-182; body
 * src/H_psi/hamiltonian_callable.cpp:94-304).  ld / ld_out in complex
 * elements (>= npw). */
int effm_hpsi(effm_ctx* ctx, int ikpt, int nb, const void* psi, int64_t ld,
              void* hpsi, int64_t ld_out, int prec, int flags);

/* dH/dk_alpha |psi>, alpha = x, y, z; a NULL output skips that direction.
 * Replaces Hamiltonian::Gradient::Callable::operator()(psi, GradHPsi)
 * (src/H_psi/hamiltonian.cppm:228-229; body
 * src/H_psi/hamiltonian_gradient.cpp:116-294). */
int effm_dhdk(effm_ctx* ctx, int ikpt, int nb, const void* psi, int64_t ld,
              void* out_x, void* out_y, void* out_z, int64_t ld_out,
              int prec, int flags);

/* Batched ragged apply: nk k-points, each with its own npw, in one set of
 * launches (up to 16 k-points per launch; every kernel walks the
 * concatenated (k-point, plane / tile, band) work of the batch).  Band-major
 * psi / hpsi blocks of nb bands per k-point at psi_ptrs[i] / hpsi_ptrs[i]
 * (device pointers), ld[i] >= npw(k) for both.  No reference counterpart
 * (the reference applies one band at one k); results equal effm_hpsi per
 * k-point. */
int effm_hpsi_multi(effm_ctx* ctx, int nk, const int* ikpts, int nb,
                    const void* const* psi_ptrs, const int64_t* ld,
                    void* const* hpsi_ptrs, int prec, int flags);

/* Batched ragged dH/dk: as effm_hpsi_multi, outputs out_x[i] / out_y[i] /
 * out_z[i] (a NULL array skips that direction for every k-point).  No
 * reference counterpart; results equal effm_dhdk per k-point. */
int effm_dhdk_multi(effm_ctx* ctx, int nk, const int* ikpts, int nb,
                    const void* const* psi_ptrs, const int64_t* ld,
                    void* const* out_x, void* const* out_y, void* const* out_z,
                    int prec, int flags);

/* H|psi> into hpsi_ptrs[i] and dH/dk_{x,y,z}|psi> into out_x/y/z[i] for the
 * same batch, as effm_hpsi_multi followed by effm_dhdk_multi with the same
 * flags (NULL direction arrays skip a direction), but with one projection
 * GEMM: the nonlocal part of H|psi> takes the <p|psi> block of the
 * gradient's <p, dp_x, dp_y, dp_z | psi> projection.  The k.p consumer needs
 * both at every k-point.  No reference counterpart; results equal the two
 * separate calls. */
int effm_hpsi_dhdk_multi(effm_ctx* ctx, int nk, const int* ikpts, int nb,
                         const void* const* psi_ptrs, const int64_t* ld,
                         void* const* hpsi_ptrs, void* const* out_x,
                         void* const* out_y, void* const* out_z, int prec,
                         int flags);

/* effm_overlap with no device allocation: the split-K partial sums live in
 * a static device array of 32 MiB (allocated with the library's code object,
 * held per GPU until the process ends) instead of a context buffer, so the
 * call makes no hipMalloc / hipFree and allocates nothing.  Results equal
 * effm_overlap's bit for bit (same split count, same kernels).  Sizes whose
 * partial sums exceed 32 MiB (an m x n block of one split: m*n*16 bytes for
 * FP64, m*n*8 for MIXED) return EFFM_ERR_INVALID.  Every call on a GPU
 * shares that GPU's array: calls that overlap in time on one GPU (other
 * streams, other threads, other contexts) race.  Device pointers only.
 * No counterpart in the reference: for callers that cannot allocate
 * (code under hipGraph capture, tight serving loops). */
int effm_overlap_noalloc(effm_ctx* ctx, int m, int n, int k, const void* A, int64_t lda,
                         const void* B, int64_t ldb, void* C, int64_t ldc, int prec, int flags);

/* z <- a z + b y + c x over n complex elements (device pointers, element
 * type from prec, arithmetic in FP64) on the context's stream: the
 * three-term recurrence of a Chebyshev filter in one pass.  No reference
 * counterpart (the reference has no eigensolver). */
int effm_lincomb(effm_ctx* ctx, int64_t n, double a, void* z, double b, const void* y,
                 double c, const void* x, int prec, int flags);

/* FP64 Rayleigh quotients E_b = Re<psi_b|hpsi_b> / <psi_b|psi_b> over npw
 * elements (test_hpsi_eigen.cpp:49-56 computes this on the host); E is a
 * device pointer unless EFFM_F_HOST.  prec selects the element type. */
int effm_rayleigh(effm_ctx* ctx, int npw, int nb, const void* psi, int64_t ld,
                  const void* hpsi, int64_t ld_h, int prec, double* E, int flags);

/* Overlap matrix C[i][j] = sum_g conj(A[i][g]) B[j][g] for i < m, j < n
 * over k coefficients (device pointers, row-major, leading dimensions in
 * complex elements): complex128 in and out for EFFM_PREC_FP64, complex64
 * for EFFM_PREC_MIXED; split-K partial sums are added in FP64 either way.
 * The overlaps / subspace Hamiltonians of an eigensolver's Rayleigh-Ritz
 * step.  No reference counterpart (the reference has no eigensolver,
 * docs/roadmap.md:5-7). */
int effm_overlap(effm_ctx* ctx, int m, int n, int k, const void* A, int64_t lda,
                 const void* B, int64_t ldb, void* C, int64_t ldc, int prec, int flags);

/* k.p effective-mass sum over states in FP64 on the device, for nsel <= 64
 * selected bands n = bands[i] (host array) of nb eigenpairs:
 *   inv_mass[i][a][b] = (d2h[i][a][b] + d2h[i][b][a]) / 2
 *       + sum_{m != n, |E_n - E_m| > degen_tol}
 *         Re(V_a[n][m] V_b[m][n] + V_b[n][m] V_a[m][n]) / (E_n - E_m),
 * i.e. the symmetrised (1/m*)_ab = <n|d2H/dk_a dk_b|n>
 * + 2 Re sum_m <n|dH_a|m><m|dH_b|n> / (E_n - E_m).  E: double[nb],
 * V[a]: complex128 [nb][ldv] row-major with V_a[m][l] = <m|dH/dk_a|l>,
 * d2h / inv_mass: double[nsel][3][3]; all device pointers.  No reference
 * counterpart: the reference declares the Hessian hooks but leaves them
 * empty (src/H_psi/hamiltonian_hessian.cpp:48-50; HessHPsi,
 * src/H_psi/hamiltonian.cppm:84-91). */
int effm_kp_mass(effm_ctx* ctx, int nb, const double* E, const void* const* V, int64_t ldv, int nsel,
                 const int* bands, const double* d2h, double degen_tol, double* inv_mass, int flags);

/* Launch shapes of one precision, the knobs the autotuner sets.  Every
 * choice gives the same results (the FFT knobs bit for bit, the GEMM ones
 * up to the order of the split-K partial sums); only the time differs. */
typedef struct effm_shapes {
    int pencil_bands; /* bands per x-pencil FFT workgroup: 4, 8, 16 */
    int pencil_z;     /* z points per x-pencil FFT workgroup: 16, 18; 0 = the
                       * grid's default (18 for the FP64 90-point x axis, else
                       * 16); effm_get_shapes reports the value in effect */
    int gemm_wgs;     /* split-K workgroup target of the projection GEMM:
                       * 512, 768, 1024, 1536, 2048 */
    int apply_tile;   /* band-tile width of the apply GEMMs above 32 bands:
                       * 64 (64 x 64 tiles), 32 (64 x 32 tiles) */
} effm_shapes;

/* Autotuner: times the candidate launch shapes with nb bands at k-point
 * ikpt in precision prec, on scratch buffers, each knob on the kernel it
 * changes (x-pencil bands and z points on the x-pencil pass, the split-K
 * target on the projection GEMM and its reduction, the tile width on the
 * apply GEMM), and keeps the fastest for every later apply in that
 * precision.  pencil_bands / gemm_wgs (may be NULL) receive those two
 * choices; effm_get_shapes returns all of them.  No reference counterpart. */
int effm_autotune(effm_ctx* ctx, int ikpt, int nb, int prec, int* pencil_bands, int* gemm_wgs);
/* The launch shapes in effect for prec (tuned, set, or the defaults
 * 8 / 0 / 1024 / 64).  No reference counterpart. */
int effm_get_shapes(const effm_ctx* ctx, int prec, effm_shapes* out);
/* Sets the launch shapes of prec (e.g. ones tuned before on the same cell);
 * a value outside its candidate list is EFFM_ERR_INVALID and changes
 * nothing.  No reference counterpart. */
int effm_set_shapes(effm_ctx* ctx, int prec, const effm_shapes* in);

/* Per-kernel timing with HIP events on the context's stream (0 = off). */
int effm_set_timing(effm_ctx* ctx, int enable);
/* Accumulated milliseconds and launch count of a kernel class; names:
 * "plane_g2r", "xpencil_vloc", "plane_r2g", "proj_gemm", "proj_reduce",
 * "proj_apply", "proj_build", "rayleigh", "overlap", "kp_sum". */
int effm_get_timing(const effm_ctx* ctx, const char* name, double* ms, int64_t* count);
int effm_reset_timing(effm_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
