/*
 * pmx_oracle.h -- C-ABI of the FP64 CPU restatement of PicoMax's EPM band and
 * chi(q,w) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that tests/, the
 * cpu_baseline leg of bench.py and __graft_entry__.smoke() compare the HIP
 * product against.  It is never linked, loaded or called by anything under
 * picomax_amd/.
 *
 * Every routine restates the reference's arithmetic loop for loop, in plain
 * C++17 + OpenMP, with its own Hermitian eigensolver (Householder + implicit
 * QL) in place of Eigen's SelfAdjointEigenSolver.  The citations on each
 * declaration name the reference function it follows.
 *
 * Complex arrays are interleaved (re, im) doubles.
 */
#ifndef PMX_ORACLE_H
#define PMX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Material + lattice description, the inputs HamiltonianEPM reads
 * (src/pmx_epm.cpp:56-153) and socCurrent reads (src/pmx_chi.cpp:203-266). */
typedef struct {
    int npw;              /* number of plane waves (global NPW)            */
    const double *G;      /* [npw][3] in 2pi/a                              */
    int natom;            /* global NATOM                                   */
    const double *tau;    /* [natom][3] atomic positions in a               */
    int nv;               /* mater::NV                                      */
    const double *g2;     /* [nv] |g|^2 keys                                */
    const double *vg;     /* [natom][nv] form factors (Ry)                  */
    double g2max;         /* mater::g2max                                   */
    int n_us;             /* symmetric SOC map (ascending keys)             */
    const double *us_key;
    const double *us_val;
    int n_ua;             /* antisymmetric SOC map (ascending keys)         */
    const double *ua_key;
    const double *ua_val;
    int use_uniform_us;   /* mater::use_uniform_us_so                       */
    double uniform_us;    /* mater::uniform_us_so_val_Ry                    */
    double a;             /* lattice constant (Angstrom), global a          */
} orc_system;

/* generate_gpt (src/pmx_basis.cpp:29-81): writes up to max_out G vectors
 * (sorted by |G|, std::sort tie order) and returns NPW, or -1. */
int orc_generate_gpt(const double *recip /*[3][3]*/, int dim, double gcut,
                     int max_out, double *G_out);

/* generate_loci (src/pmx_basis.cpp:142-179), brute-force search. */
int orc_generate_loci(int npw, const double *G, int neps, int *loci /*[neps][npw]*/);

/* HamiltonianEPM (src/pmx_epm.cpp:56-153): full 2npw x 2npw matrix,
 * column-major, lower triangle filled as the reference fills it and the
 * upper triangle set to the Hermitian conjugate. */
int orc_hamiltonian(const orc_system *sys, const double *K, double *H);

/* Hermitian eigensolver (role of SelfAdjointEigenSolver, Eigen's
 * SelfAdjointEigenSolver.h:420-467): reads the lower triangle of the n x n
 * column-major complex A, returns ascending eigenvalues w and, when Z is not
 * NULL, orthonormal eigenvectors as columns of Z. */
int orc_eigh(int n, const double *A, double *w, double *Z);

/* setRefEnergy (src/pmx_epm.cpp:393-428). */
int orc_ref_energy(const orc_system *sys, const double *Kref, int nvalence,
                   double *offset);

/* empiricalpseudopotentialmethod's loop (src/pmx_epm.cpp:225-248):
 * eband[q][n] = e_n(Q_q) - offset for n < nband (0 beyond 2npw). */
int orc_bands(const orc_system *sys, int nq, const double *Q, int nband,
              double offset, double *eband);

/* chi_tensor (src/pmx_chi.cpp:319-713) including the optional KK pass.
 * Output ReX/ImX in the file order of export_Xij (src/pmx_io.cpp:720-741):
 * [i][j][q][m][n<=m][f]. */
int orc_chi_tensor(const orc_system *sys, const int *loci, int neps,
                   int nk, const double *K, const double *KW,
                   int nq, const double *Q, int nband, double offset,
                   int delta, double epsilon, int nfreq, double dfreq,
                   int kk, double bz_volume, double *ReX, double *ImX);

/* chi_LL in the spinor basis of the current sources (src/pmx_chi_copy.cpp:
 * 320-676; src_bkp/pmx_chi.cpp:126-439 is the same sum over a spatial basis):
 * density vertex for i = 0, no frequency window, scale factors LL/LT/TT, the
 * principal-value term when kk == 0, f = 0 of the delta term zeroed.  Same
 * arguments and output layout as orc_chi_tensor. */
int orc_chi_ll(const orc_system *sys, const int *loci, int neps,
               int nk, const double *K, const double *KW,
               int nq, const double *Q, int nband, double offset,
               int delta, double epsilon, int nfreq, double dfreq,
               int kk, double bz_volume, double *ReX, double *ImX);

/* chi_xyz, the -switch epsxyz tensor (src_bkp/pmx_chi.cpp:723-1000 in the
 * spinor basis: same-spin products): bra = valence at k, ket = conduction at
 * k+q, vertex x_i.(G_p + K + Q/2 + G_m/2), no window, SCALEFACTOR times
 * (pi delta-sum / w^2 + 2 principal value with 1/dE/(dE^2 - w^2)). */
int orc_chi_xyz(const orc_system *sys, const int *loci, int neps,
                int nk, const double *K, const double *KW,
                int nq, const double *Q, int nband, double offset,
                int delta, double epsilon, int nfreq, double dfreq,
                int kk, double bz_volume, double *ReX, double *ImX);

/* kramerskronigtransform (src/pmx_chi.cpp:272-309) on one element. */
void orc_kk(int nfreq, double dfreq, const double *w, double *ReX, double *ImX);

/* Broadening functions (src/pmx_chi.cpp:90-164); delta codes as in
 * chi_tensor's switch (src/pmx_chi.cpp:324-337): 0 Gaussian, 1 Lorentzian,
 * 2 triangle, 3 rect. */
double orc_delta(int delta, double epsilon, double x);

/* Window half-width (src/pmx_chi.cpp:564-570). */
int orc_window_half(int delta, double epsilon, double dfreq);

/* uvec_LT (src/pmx_basis.cpp:830-874): 3 complex 3-vectors [3][3][2]. */
void orc_uvec_lt(const double *v, double *u);

/* Last error message of this thread. */
const char *orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
