/*
 * pmx.h -- C-ABI of libpmx, the MI355X (gfx950) implementation of PicoMax's
 * EPM band-structure and chi(q,w) path.
 *
 * The reference has no library interface: its CLI calls C++ functions on
 * global state.  Each entry point below replaces one of those functions (the
 * citation names it); the `picomax` CLI, the Python wrapper and the tests all
 * go through this header.
 *
 * Conventions
 *   - every function returns 0 on success or a negative PMX_E* code and never
 *     aborts; pmx_last_error() gives the message (thread-local);
 *   - arrays are plain host pointers owned by the caller; complex arrays are
 *     interleaved (re, im) doubles; lengths are given explicitly;
 *   - a pmx_ctx owns one HIP device's buffers and stream and is used from
 *     one host thread at a time; calls are synchronous on return;
 *   - output layouts equal the reference's file layouts, so the CLI writes
 *     them verbatim.
 */
#ifndef PMX_H
#define PMX_H

#ifdef __cplusplus
extern "C" {
#endif

#define PMX_OK 0
#define PMX_EINVAL (-1)   /* bad argument or inconsistent sizes            */
#define PMX_EHIP (-2)     /* HIP runtime error                               */
#define PMX_ENOMEM (-3)   /* device allocation failed                        */
#define PMX_ECONV (-4)    /* eigensolver did not converge                    */
#define PMX_ESTATE (-5)   /* call order (no system set, ...)                 */
#define PMX_EIO (-6)      /* file could not be read or written               */
#define PMX_EBANDS (-7)   /* no sign change of E within -nband (see pmx_chi) */

/* precision of the chi path (the eigensolve is FP64 in every mode) */
#define PMX_PREC_FP64 0    /* FP64 throughout                                */
#define PMX_PREC_FP32 1    /* rho + rank update FP32 MFMA, FP64 k-reduction  */
#define PMX_PREC_BF16X3 2  /* split-BF16 (hi+lo, 3 products) MFMA, FP32 acc  */
#define PMX_PREC_DIRECT 3  /* FP64, per-transition windowed sum (no binning) */

typedef struct pmx_ctx pmx_ctx;

/* Everything HamiltonianEPM reads (src/pmx_epm.cpp:56-153) and chi_tensor
 * reads besides the k/q sets (src/pmx_chi.cpp:319-713). */
typedef struct {
    int npw;               /* NPW                                            */
    const double *G;       /* [npw][3] G vectors in 2pi/a, generate_gpt order */
    int neps;              /* NEPS local-field G (the first neps of G)       */
    int natom;             /* NATOM                                          */
    const double *tau;     /* [natom][3] atomic basis in a                   */
    int nv;                /* number of form-factor shells                   */
    const double *g2;      /* [nv] |G|^2 keys (2pi/a)^2                      */
    const double *vg;      /* [natom][nv] form factors, Ry                   */
    double g2max;          /* g2[nv-1] + 1e-2 (init_mater)                   */
    int n_us;              /* symmetric SOC map, ascending keys              */
    const double *us_key;
    const double *us_val;
    int n_ua;              /* antisymmetric SOC map, ascending keys          */
    const double *ua_key;
    const double *ua_val;
    int use_uniform_us;    /* -use_uniform_us_so                             */
    double uniform_us;     /* -uniform_us_so_val, Ry                         */
    double a;              /* lattice constant, Angstrom                     */
    double bz_volume;      /* lattice::bz_volume, (2pi/a)^3                  */
} pmx_system;

/* Inputs of chi_tensor (src/pmx_chi.cpp:319-713). */
typedef struct {
    int nk;                /* NKPT                                           */
    const double *K;       /* [nk][3] k-points, 2pi/a                        */
    const double *KW;      /* [nk] weights (already divided by NKPT0)        */
    int nq;                /* NQ                                             */
    const double *Q;       /* [nq][3] q-vectors, 2pi/a                       */
    int nband;             /* NBAND (spinor count)                           */
    double offset;         /* env::energyoffset, eV                          */
    int delta;             /* 0 Gaussian, 1 Lorentzian, 2 triangle, 3 rect   */
    double epsilon;        /* EPSILON, eV                                    */
    int nfreq;             /* NFREQ; w_f = f * dfreq                         */
    double dfreq;          /* eV                                             */
    int kk;                /* apply kramerskronigtransform                   */
    int precision;         /* PMX_PREC_*                                     */
} pmx_chi_params;

/* ---------------------------------------------------------------- context */
int pmx_create(int device, pmx_ctx **out);
void pmx_destroy(pmx_ctx *ctx);
const char *pmx_last_error(void);
/* library version string */
const char *pmx_version(void);

/* Uploads the system, builds the plane-wave index map loci on the device
 * (replaces generate_loci, src/pmx_basis.cpp:142-179) and the k-independent
 * form-factor / SOC pair tables used by H(k) (HamiltonianEPM,
 * src/pmx_epm.cpp:56-153; pseudopotential, :171-203; get_soc_form_factor,
 * :22-50). */
int pmx_set_system(pmx_ctx *ctx, const pmx_system *sys);

/* Replaces only the form factors vg [natom][nv] (Ry) of the current system. */
int pmx_set_form_factors(pmx_ctx *ctx, const double *vg);

/* Copies the device loci table [neps][npw] (index or -1). */
int pmx_get_loci(pmx_ctx *ctx, int *loci);

/* setRefEnergy (src/pmx_epm.cpp:393-428): offset = (e[nv-1] + e[nv]) / 2 at
 * Kref, or 0 when nvalence is out of range. */
int pmx_ref_energy(pmx_ctx *ctx, const double *Kref, int nvalence, double *offset);

/* The band loop of empiricalpseudopotentialmethod (src/pmx_epm.cpp:225-248):
 * eband[q][n] = e_n(Q_q) - offset for n < nband, 0 for n >= 2*npw. */
int pmx_bands(pmx_ctx *ctx, int nq, const double *Q, int nband, double offset, double *eband);

/* Batched FP64 Hermitian eigensolve of H(k) (HamiltonianEPM +
 * SelfAdjointEigenSolver, src/pmx_epm.cpp:232; src/pmx_chi.cpp:388, 476):
 * lowest nband spinor eigenvalues evals [nk][nband] and, when evecs is not
 * NULL, their eigenvectors evecs [nk][nband][2*npw] complex, spinor rows
 * interleaved (2p = up, 2p+1 = down). */
int pmx_eig_batch(pmx_ctx *ctx, int nk, const double *K, int nband, double *evals, double *evecs);

/* pmx_eig_batch for several systems in one launch (no reference counterpart):
 * batch i solves nk[i] k-points K[i] of context ctxs[i]'s system, SOC or not,
 * any basis size, into evals[i] [nk[i]][nband[i]] and, when evecs and
 * evecs[i] are not NULL, evecs[i] [nk[i]][nband[i]][2*npw_i] (the layouts of
 * pmx_eig_batch).  The contexts must be distinct and on one device.  The
 * batches run side by side, each on its own context's stream, forked from
 * and joined to the first context's stream; with graphs on (pmx_graphs, kept
 * by the first context) the second call with the same shapes is captured as
 * one HIP graph with a branch per batch, and later calls replay it. */
int pmx_eig_batch_multi(int nsys, pmx_ctx *const *ctxs, const int *nk, const double *const *K, const int *nband,
                        double *const *evals, double *const *evecs);

/* chi_tensor (src/pmx_chi.cpp:319-713) including the optional KK pass.
 * ReX and ImX each hold 9*nq*neps*(neps+1)/2*nfreq doubles in the order
 * [i][j][q][m][n<=m][f] of export_Xij (src/pmx_io.cpp:720-741). */
int pmx_chi(pmx_ctx *ctx, const pmx_chi_params *p, double *ReX, double *ImX);

/* pmx_chi with ReX / ImX in device memory (hipMalloc'd by the caller, same
 * layout and sizes): nothing crosses PCIe but the k / q lists. */
int pmx_chi_device(pmx_ctx *ctx, const pmx_chi_params *p, double *dReX, double *dImX);

/* chi_LL, the -switch epsLL tensor (src/pmx_chi_copy.cpp:320-676 in the
 * spinor basis of the current sources; src_bkp/pmx_chi.cpp:126-439 is the
 * same sum over a spatial basis): density vertex for i = 0, every transition
 * at every frequency (no window), LL/LT/TT scale factors over |q+G_m|,
 * |q+G_n| and w, the delta term zeroed at f = 0, and with kk == 0 the
 * principal-value term (per transition, FP64).  Same parameters and output
 * layout as pmx_chi; the _device form takes device output pointers. */
int pmx_chi_ll(pmx_ctx *ctx, const pmx_chi_params *p, double *ReX, double *ImX);
int pmx_chi_ll_device(pmx_ctx *ctx, const pmx_chi_params *p, double *dReX, double *dImX);

/* chi_xyz, the -switch epsxyz tensor (src_bkp/pmx_chi.cpp:723-1000 with
 * dispatch src_bkp/pmx.cpp:226-244; src/pmx_chi_copy.cpp:1070-1344 is the same
 * loop but copies only NPW components of its 2 NPW spinor eigenvectors):
 * bra = valence state at k, ket = conduction state at k+q, vertex
 * x_i.(G_p + K + Q/2 + G_m/2) summed over equal spins, every transition at
 * every frequency, SCALEFACTOR (the TT one of chi_LL) times (pi delta-sum/w^2,
 * zero at f = 0, + 2 principal value with 1/dE/(dE^2 - w^2) when kk == 0).
 * Same parameters and output layout as pmx_chi. */
int pmx_chi_xyz(pmx_ctx *ctx, const pmx_chi_params *p, double *ReX, double *ImX);
int pmx_chi_xyz_device(pmx_ctx *ctx, const pmx_chi_params *p, double *dReX, double *dImX);

/* Per-kernel timing with HIP events on the context's stream (no reference
 * counterpart).  enable != 0 starts recording (and clears), 0 stops.
 * pmx_profile_report writes a JSON object
 * {"kernel": {"count": n, "ms": total, "flops": f, "bytes": b}, ...}
 * into buf (NUL-terminated, truncated to len). */
int pmx_profile(pmx_ctx *ctx, int enable);
int pmx_profile_report(pmx_ctx *ctx, char *buf, int len);

/* Eigensolve batches as HIP graphs (no reference counterpart).  With graphs
 * on (the default), the first eigensolve batch of a shape runs eagerly, the
 * second identical one (the same sizes, system and buffer addresses) is
 * captured into a graph, and later ones replay it: the k+q solves of every q
 * after the second in a chi q list, and repeated pmx_eig_batch calls.
 * Profiling (pmx_profile) and PMX_GRAPHS=0 keep every batch eager.
 * enable: 1 on, 0 off (drops the graphs held), -1 leaves the setting.
 * stats (may be NULL): {graphs held, captures, replays} so far. */
int pmx_graphs(pmx_ctx *ctx, int enable, long *stats);

/* Whole q steps of a chi q list as HIP graphs (no reference counterpart; the
 * loop it replaces is src/pmx_chi.cpp:462-658).  With graphs on and
 * profiling off, a pmx_chi / pmx_chi_device call over three or more q (the
 * binned moment path) runs its first two q steps eagerly and captures the
 * third — k+q eigensolve, classification, binning, rho, rank update,
 * convolution, KK — into one graph that reads its q from device memory;
 * every later q replays it with no host readback.  A replayed step whose
 * conduction / valence counts differ from the captured ones is run again
 * eagerly.  stats: {captures, replays, steps run again} so far. */
int pmx_qgraph_stats(pmx_ctx *ctx, long *stats);

/* The autotuner's decisions so far on this context, as a JSON list: per
 * kernel shape the candidates' times on the sample it ran and the variant it
 * keeps (no reference counterpart). */
int pmx_tune_report(pmx_ctx *ctx, char *buf, int len);

/* kramerskronigtransform (src/pmx_chi.cpp:272-309) applied in place to nel
 * contiguous elements of nfreq points each (w_f = f * dfreq). */
int pmx_kk(pmx_ctx *ctx, int nel, int nfreq, double dfreq, double *ReX, double *ImX);

/* Local-field inversion (new stage; the reference only has it in MATLAB,
 * include/matlab_wrapper/+picomax/+eps/macroscopic.m:23-38, 87-89):
 * eps_mn = delta_mn + chi^LL_mn (i = j = 0 block, lower triangle + Hermitian
 * fill), per (q, f): epsM = 1 / [eps^-1]_00, elf = -Im(1/epsM).
 * ReX/ImX as returned by pmx_chi; epsM [nq][nfreq] complex, elf [nq][nfreq];
 * epsinv [nq][nfreq][neps][neps] complex may be NULL. */
int pmx_epsinv(pmx_ctx *ctx, int nq, int neps, int nfreq, const double *ReX, const double *ImX,
               double *epsM, double *elf, double *epsinv);

/* pmx_epsinv with no device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  dReX / dImX are
 * device arrays whose leading [nq][neps(neps+1)/2][nfreq] doubles are the
 * i = j = 0 block (pmx_chi_device's output qualifies); dEpsM [nq][nfreq]
 * complex and dElf [nq][nfreq] are device outputs, as pmx_epsinv's epsM /
 * elf.  The Gauss-Jordan workspace is a fixed 32 MiB device array per GPU,
 * set up by pmx_create; neps above 1024 (more than 32 MiB for one matrix) is
 * rejected with PMX_EINVAL.  All calls on one GPU share that workspace:
 * calls that overlap on one GPU (several host threads, several contexts)
 * race.  Synchronous on the context's stream. */
int pmx_epsinv_noalloc(pmx_ctx *ctx, int nq, int neps, int nfreq, const double *dReX, const double *dImX,
                       double *dEpsM, double *dElf);

/* The MATLAB wrapper's macroscopic dielectric functions
 * (include/matlab_wrapper/+picomax/+eps/macroscopic.m:1-89), from all nine
 * blocks of a chi_LL result (pmx_chi_ll's ReX/ImX for the same nq, Q, neps,
 * nfreq, dfreq; Q [nq][3] in 2 pi / a units; G are the context's first neps
 * plane waves, the .gpt rows).  eps_L = I + chi[L][L], eps_T = I + chi[T+][T+]
 * (lower triangle, conjugate fill); with nq = 1 and Q = 0 the head of eps_L is
 * eps_T's (:18-20).  eps_T's diagonal gets -c0^2 |q + G_m|^2 / w^2 with the
 * MATLAB's constants (:51-53, 70-78): c0 = 3e8 m/s, w = omega / 4.135667696e-15
 * (eV s), |q + G| = 2 pi |Q + G| / (a * 1e-10 m), and |q + G|^2 = 0 taken as
 * 1e-15.  The MATLAB takes G from a fixed list of nine fcc vectors (:41-50);
 * here they are the system's own.
 * Outputs, complex [nq][nfreq]: eps00 = eps_L[0][0], epsM = 1/[eps_L^-1]_00,
 * epsM_T = 1/[eps_T^-1]_00 (NaN at f = 0, where w = 0). */
int pmx_eps_macro(pmx_ctx *ctx, int nq, const double *Q, int neps, int nfreq, double dfreq,
                  const double *ReX, const double *ImX, double *eps00, double *epsM, double *epsM_T);

/* The MATLAB wrapper's polarization bands
 * (include/matlab_wrapper/@picomax/polarizationband.m:1-121): per (q, f) the
 * eigenvalues of eps = I + chi[blk][blk] (blk 0 = L, 1 = T+, 2 = T-; stored
 * lower triangle, conjugate fill above, complex diagonal as stored, so a
 * general complex eigenproblem), tracked across f as eigenshuffle does
 * (include/matlab_wrapper/+picomax/+util/eigenshuffle.m): at f = 0 the
 * first q in descending real part and each next q matched to the one before
 * (:12-30, 58-66), then along f each frequency matched to the previous one
 * (:33-57), a match being the minimum-cost assignment of
 * (1 - |v_i^H v_j|) |d_i - d_j|.  Where an f = 0 eigenvalue is repeated the
 * MATLAB's 1e-10 lookup (:62) copies one track over its partner; here both
 * tracks are kept.  ReX/ImX: all
 * nine blocks as pmx_chi/pmx_chi_ll return them.  alpha [nq][nfreq][neps]
 * complex; vec [nq][nfreq][neps][neps] complex (vector s of slot s, unit
 * norm, largest component real, eigenshuffle's sign rule along f) may be NULL.
 * PMX_EINVAL if an eigenvalue iteration does not converge. */
int pmx_polarization_bands(pmx_ctx *ctx, int nq, int neps, int nfreq, int blk, const double *ReX,
                           const double *ImX, double *alpha, double *vec);

/* Hellmann-Feynman adjoint of the band energies (no reference counterpart):
 * given upstream gradients gE [nk][nband] of eband (band energies before the
 * offset), returns gvg [natom][nv] = sum_k,n gE[k][n] dE_n(k)/dvg. */
int pmx_grad_bands(pmx_ctx *ctx, int nk, const double *K, int nband, const double *gE, double *gvg);

/* Adjoint of chi_tensor (src/pmx_chi.cpp:319-713) with respect to the form
 * factors (no reference counterpart): given upstream gradients gRe, gIm of
 * the ReX/ImX that pmx_chi returns for the same p (same layout), returns
 * gvg [natom][nv] = d/dvg sum (gRe ReX + gIm ImX) of the exact windowed sum
 * in FP64, whatever p->precision says.  The energy zero p->offset is held
 * fixed (it only classifies states).  Gaussian or Lorentzian broadening;
 * nband even for the spatial solve (a Kramers pair is one solve state), any
 * value up to 2 npw with spin-orbit maps; PMX_EINVAL otherwise.  Eigenvector
 * derivatives by first-order perturbation over the whole basis (N = npw, or
 * 2 npw spinor components with SOC), pairs closer than 1e-8 eV masked.  The
 * k-points go in batches of at most 512 (fewer for large N, about 4 GB of
 * eigenvectors at k and k+q per batch) whose eigenpairs are recomputed there,
 * so device memory does not grow with the grid. */
int pmx_grad_chi(pmx_ctx *ctx, const pmx_chi_params *p, const double *gRe, const double *gIm, double *gvg);

/* --------------------------------------------------------- host set-up
 * The reference's CLI set-up (inputParser src/pmx.hpp:190-245, inputProcess
 * src/pmx_io.cpp:423-503, init_crystal :108-310, init_mater :315-417,
 * generate_gpt src/pmx_basis.cpp:29-81, import_kpt :200-245) as a parsed,
 * opaque input object.  Host only; needs no GPU. */
typedef struct pmx_input pmx_input;

int pmx_input_parse(int argc, const char *const *argv, pmx_input **out);
void pmx_input_free(pmx_input *in);
/* pointers into the input's own storage, valid until pmx_input_free */
int pmx_input_system(const pmx_input *in, pmx_system *out);

typedef struct {
    const char *sw;        /* -switch value ("" if absent)                   */
    const char *outputfile;
    const char *wdir;
    int debug;
    int nband, nvalence, neps, kk, delta, nfreq, npw, natom, nq, nk, kpoint;
    double epsilon, dfreq, gcut, encut, a, bz_volume;
    double refpoint[3];
    const double *Q;       /* [nq][3]                                        */
    const double *K;       /* [nk][3] (NULL when there are no k-points)      */
    const double *KW;      /* [nk]                                           */
    int kpointorder;       /* -kpointorder of the -kpoint 1/2/3 generators   */
} pmx_run_info;

int pmx_input_info(const pmx_input *in, pmx_run_info *out);

/* The reference main's set-up report on stdout (src/pmx.cpp:84-160): banner,
 * host, parameters, the Q path, the G count and the k-point source with
 * their set-up times; with -debug > 1 also printInput (src/pmx_io.cpp:510-614:
 * options, form factors, SOC maps, atomic basis, and with -debug > 2 every G,
 * K and Q vector in Eigen's row format).  precision != PMX_PREC_FP64 adds one
 * "precision" line to the parameters (the option the reference does not have). */
int pmx_print_setup(const pmx_input *in, int precision);

/* generate_gpt on its own: up to max_out G vectors sorted like the reference;
 * returns NPW (>= 0) or an error code. */
int pmx_generate_gpt(const double *recip, int dim, double gcut, int max_out, double *G_out);

/* The CLI body after parsing (src/pmx.cpp:171-203): runs -switch band |
 * epsij on ctx, and epsLL / epsxyz as the older snapshot's main dispatches
 * them (src_bkp/pmx.cpp:206-244; the current main has no such branches),
 * prints the reference's solver progress lines to stdout (chi_tensor's
 * per-phase and per-q lines with their times, src/pmx_chi.cpp:370-657) and
 * writes ${wdir}/${outputfile}.dat (+ .gpt for the chi switches).  An absent
 * or unknown -switch does nothing and returns 0.  precision selects the chi
 * path (PMX_PREC_*). */
int pmx_run(pmx_ctx *ctx, const pmx_input *in, int precision);

#ifdef __cplusplus
}
#endif
#endif
