/*
 * effm_oracle.h — FP64 CPU restatement of the effmass H|psi> hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle: it is linked or
 * loaded only by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg.  The product path (libeffm_hip.so) never calls it.
 *
 * Every function follows a reference file/line range (paths relative to the
 * reference repository root) and keeps its loop structure, so that timing it
 * is a fair CPU baseline for the reference's single-threaded code.
 */
#ifndef EFFM_ORACLE_H
#define EFFM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One pseudopotential species after sortByL + diagonalizeNonlocal
 * (src/pseudo/ncpp_advance.cppm:11-153): nbeta radial channels sorted by l,
 * lambda = diagonal of the rotated D block, raw beta(q) / dbeta/dq nodes on
 * q_i = i*dq (src/pseudo/ncpp_advance.cppm:586-641, 781-835). */
typedef struct {
    int lmax;               /* -1: no projectors */
    int nbeta;
    const int* beta_l;      /* [nbeta], non-decreasing */
    const double* lambda;   /* [nbeta] Hartree */
    int nq;
    double dq;
    double qmax;
    const double* beta_tab;   /* [nbeta*nq] raw nodes (no 1/sqrt(Omega)) */
    const double* dbeta_tab;  /* [nbeta*nq] or NULL */
} orc_species;

typedef struct {
    int n1, n2, n3;
    const double* vloc;     /* [n1*n2*n3] Hartree, index i*n2*n3 + j*n3 + k */
    double omega;           /* bohr^3 */
    int natom;
    const int* species;     /* [natom] species index, atoms grouped by species */
    const double* tau;      /* [natom*3] fractional coordinates */
    int nspecies;
    const orc_species* sp;  /* [nspecies] in ascending-Z order */
} orc_cell;

typedef struct {
    int npw;
    const double* K;        /* [npw*3] Cartesian K = k+G (bohr^-1) */
    const double* kin;      /* [npw] 0.5|K|^2 Hartree */
    const int32_t* gidx;    /* [npw*3] centred integer G components */
    double kfrac[3];        /* fractional k */
} orc_kpt;

/* 1D / 3D FFT (src/math/fft.cppm:18-93): forward = exp(-i), scaled 1/n per
 * axis; backward = exp(+i), unscaled.  data is interleaved re,im. */
void orc_fft1d(double* data, int n, int forward);
void orc_fft3d(double* grid, int n1, int n2, int n3, int forward);

/* beta / dbeta interpolation (src/pseudo/ncpp_advance.cppm:644-674, 838-868). */
int orc_interp(const double* tab, int nq, double dq, double qmax,
               double norm, double q, double* out);

/* Real Y_lm and angular gradients for one direction
 * (src/math/sph_harmonics.cppm:120-335, 580-875).  out arrays are indexed
 * l*l + (m+l) for l = 0..lmax. */
void orc_ylm(double theta, double phi, int lmax, double* ylm,
             double* dtheta, double* dphi_over_sin);

/* Separable structure factor (src/H_psi/structure_factor.cppm:92-159). */
void orc_sfac(const double tau[3], const int g[3], const double kfrac[3],
              int n1, int n2, int n3, double out[2]);

/* H|psi> for nb bands (psi, hpsi: [nb][npw] interleaved complex, ld = npw),
 * accumulate semantics (src/H_psi/hamiltonian_callable.cpp:94-304).
 * Returns 0 on success, negative on an input error (q out of table range). */
int orc_hpsi(const orc_cell* cell, const orc_kpt* kp, int nb,
             const double* psi, double* hpsi);

/* Only the nonlocal part of H|psi> (hamiltonian_callable.cpp:200-304). */
int orc_vnl(const orc_cell* cell, const orc_kpt* kp, int nb,
            const double* psi, double* hpsi);

/* dH/dk_alpha |psi>, alpha = x,y,z; NULL output = skip direction
 * (src/H_psi/hamiltonian_gradient.cpp:116-294).  Accumulate semantics. */
int orc_dhdk(const orc_cell* cell, const orc_kpt* kp, int nb,
             const double* psi, double* ox, double* oy, double* oz);

/* Band-parallel H|psi> over nthreads OpenMP threads (one work set per
 * thread); used only as the all-core CPU baseline. */
int orc_hpsi_parallel(const orc_cell* cell, const orc_kpt* kp, int nb,
                      const double* psi, double* hpsi, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
