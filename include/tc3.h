/*
 * tc3 -- MI355X three-phonon scattering and lattice thermal conductivity engine.
 *
 * C ABI of libtc3.so.  Plain pointers and sizes, no C++ or torch types.  All
 * quantities are in the reference's internal Hartree atomic units (lengths in
 * Bohr, reciprocal vectors in 1/Bohr without 2*pi, masses in electron masses,
 * IFCs in Hartree/Bohr^n, frequencies in Hartree).  Fortran column-major
 * conventions are kept for 3-vectors and 3x3 matrices: v[3*i + c] is
 * component c of vector i, and m[3*j + i] is element (i, j).
 *
 * Every function returns 0 on success or a positive status that follows the
 * reference's exit codes (src/libolle/konstanter.f90:160-172):
 *   1 bad dimension, 2 eigen-solver failure, 3 unphysical input,
 *   4 symmetry failure, 5 bad parameter, 6 I/O, 7 device error.
 * tc3_last_error() returns the message of the last failure on a context.
 * No C++ exception crosses this boundary.  One context per host thread; calls
 * on one context are ordered and not re-entrant.  Host arrays belong to the
 * caller and are copied in; device memory belongs to the context and is
 * released by tc3_destroy.  One GPU, no collectives.
 */
#ifndef TC3_H
#define TC3_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tc3_ctx tc3_ctx;

enum {
    TC3_PREC_FP64 = 0,   /* |Psi|^2 contraction in FP64                       */
    TC3_PREC_FP32 = 1,   /* FP32 MFMA, FP32 accumulation                       */
    TC3_PREC_BF16X3 = 2  /* split-BF16 x3 MFMA, FP32 accumulation              */
};

/* Create / destroy a context bound to one HIP device.  device = -1 creates a
 * host-only context that accepts the structure, IFCs and q-mesh calls and refuses
 * every device computation.  No counterpart in the reference (its state lives in
 * Fortran derived types owned by main.f90). */
int tc3_create(tc3_ctx **ctx, int device);
void tc3_destroy(tc3_ctx *ctx);
const char *tc3_last_error(const tc3_ctx *ctx);

/* Crystal structure: replaces lo_crystalstructure%readfromfile + readisotopefromfile
 * (type_crystalstructure_io.f90:9-350, :921-959).  latvec: 3x3 lattice vectors as
 * columns (Bohr); rfrac: 3*na fractional positions; species: na species ids (atoms
 * with equal ids are interchangeable by symmetry); mass: na masses (emu). */
int tc3_set_structure(tc3_ctx *ctx, int na, const double *latvec, const double *rfrac,
                      const int *species, const double *mass);

/* Natural isotope composition of an element: replaces lo_isotope_distribution%naturaldistribution
 * (type_crystalstructure.f90:1521-2466), which gives every atom its masses when infile.isotopes
 * is not read (--readiso off, the reference's default; type_crystalstructure.f90:650-652).
 * symbol: element symbol ("Si"); *n receives the number of isotopes; conc and mass (amu)
 * receive the abundances and masses when maxn >= *n.  The caller normalises and averages them
 * as for infile.isotopes (type_crystalstructure_io.f90:953-957).  Returns 5 for an unknown
 * symbol, 1 when maxn is too small for non-NULL arrays.  Needs no context. */
int tc3_natural_isotopes(const char *symbol, int maxn, double *conc, double *mass, int *n);

/* Second-order IFCs: replaces lo_forceconstant_secondorder%readfromfile
 * (type_forceconstant_secondorder_io.f90:109-247), non-polar.  a1, a2: 0-based atoms;
 * lv2: 3*npair Cartesian lattice vectors (Bohr); m: 9*npair tensors (Ha/Bohr^2),
 * m[9*n + 3*j + i] = Phi_ij. */
int tc3_set_fc2(tc3_ctx *ctx, int npair, const int *a1, const int *a2, const double *lv2,
                const double *m);

/* Third-order IFCs: replaces lo_forceconstant_thirdorder%readfromfile
 * (type_forceconstant_thirdorder.f90:243-311).  lv2, lv3: 3*ntrip Cartesian lattice
 * vectors (Bohr); m: 27*ntrip tensors (Ha/Bohr^3), m[27*t + 9*k + 3*j + i] = Phi_ijk.
 * The mass weighting (mwm, :277) is applied inside. */
int tc3_set_fc3(tc3_ctx *ctx, int ntrip, const int *a1, const int *a2, const int *a3,
                const double *lv2, const double *lv3, const double *m);

/* Isotope scattering (compute_isotope_scattering, scattering_isotope.f90:2-92): disorder
 * holds the mass-disorder parameter g of every atom (mass_disorder_parameter,
 * type_crystalstructure.f90:1505-1519, from infile.isotopes); enable = 1 adds the isotope
 * term to the linewidths and the scattering matrix of every later tc3_scatter call (the
 * reference's default; its --noisotope corresponds to enable = 0, the state after
 * tc3_set_structure).  The term is zero when every g is zero. */
int tc3_set_isotopes(tc3_ctx *ctx, const double *disorder, int enable);

/* Boundary scattering: a mean-free-path limit L (Bohr; the reference's --max_mfp in metres,
 * converted at options.f90:218) adds |v|/L to the linewidth of every mode whose group velocity
 * exceeds 1e-10 a.u., after the isotope and three-phonon terms and before the degeneracy
 * average (scattering.f90:245-250).  L <= 0 switches it off (the default).  Replaces the
 * opts%mfp_max input of lo_scattering_rates%generate. */
int tc3_set_boundary(tc3_ctx *ctx, double mfp_max);

/* Monte-Carlo integration grid of the three-phonon term (-qg3ph, --seed; options.f90:116,
 * 226-230; lo_montecarlo_grid, lo_fftgrid_helper.f90:30-99): for every row (q1, b1) a
 * stratified random draw of prod(qg3ph) points of the q-mesh from the reference's 64-bit
 * Mersenne twister seeded as lo_mersennetwister%init(0, 1/seed) (one MPI rank), each draw
 * weighted 1/prod(qg3ph).  qg3ph NULL, or every entry outside 1..qgrid, means the full
 * mesh; a Monte-Carlo grid needs seed > 0 (the reference's wall-clock seed is not
 * reproducible).  Set after tc3_set_qmesh. */
int tc3_set_montecarlo(tc3_ctx *ctx, const int *qg3ph, int seed);

/* Space group, Gamma-centred FFT q-mesh, irreducible points, little groups and the
 * irreducible triplet lists: replaces uc%classify('wedge') (type_crystalstructure_symmetry.f90:7-91),
 * lo_generate_qmesh (type_qpointmesh.f90:452-657, type_qpointmesh_gridgeneration.f90:8-545)
 * and triplet_is_irreducible (scattering_threephonon.f90:319-448).  Host-side.  Resets the
 * three- and four-phonon Monte-Carlo grids to the full mesh (the defaults of -qg3ph and
 * -qg4ph, options.f90:226-236). */
int tc3_set_qmesh(tc3_ctx *ctx, const int *dims, int timereversal);

/* tc3_set_qmesh with the reference's hidden --nosym (options.f90:157, 206; lo_generate_qmesh
 * nosym, type_qpointmesh_gridgeneration.f90:54-58): qpsymmetry = 0 reduces the mesh by the
 * identity and time reversal only (every point its own little group), while the thermal
 * conductivity is still averaged over the crystal's space group (kappa.f90:28-75, 308-324).
 * tc3_set_qmesh is qpsymmetry = 1. */
int tc3_set_qmesh_sym(tc3_ctx *ctx, const int *dims, int timereversal, int qpsymmetry);

/* Three-phonon term on (enable = 1, the default) or off (the reference's --nothirdorder,
 * options.f90:77-80, 187-188; scattering.f90:119-121, 234-238): without it tc3_scatter forms
 * the linewidths and scattering matrix from the isotope, four-phonon and boundary terms only
 * (and fails with status 5 when none is on), no three-phonon Monte-Carlo grid is drawn, and
 * the adjoints return status 5. */
int tc3_set_thirdorder(tc3_ctx *ctx, int enable);

/* Sizes of what tc3_set_qmesh built. */
int tc3_mesh_info(const tc3_ctx *ctx, int *nsym, int *nfull, int *nirr, int64_t *ntriplets,
                  int *nrows);

/* Mesh arrays: irreducible full indices (nirr), irreducible index of every full point
 * (nfull), operation from irreducible (nfull, 1-based, negative = time reversal),
 * integration weights (nirr), Cartesian q of every full point (3*nfull). */
int tc3_get_mesh(const tc3_ctx *ctx, int *irr_full_index, int *irr_of_full, int *op_of_full,
                 double *weight_irr, double *qcart);

/* Symmetry operations: Cartesian rotations (9*nsym, column-major), Cartesian
 * translations (3*nsym), atom maps (na*nsym). */
int tc3_get_symmetry(const tc3_ctx *ctx, double *rot, double *tr, int *fmap);

/* Harmonic properties on the mesh on the GPU: dynamical matrix, FP64 Hermitian
 * eigen-solve, degenerate-subspace rotation, group velocities, unfolding to the full
 * grid.  Replaces lo_phonon_dispersions%generate (type_phonon_dispersions_generation.f90:8-241)
 * with fc%dynamicalmatrix / frequencies_eigenvectors_groupvelocities
 * (type_forceconstant_secondorder_dynamicalmatrix.f90:13-481, :484-788).
 * omega_min receives the smallest non-zero frequency (unstable if < -1e-9). */
int tc3_harmonic(tc3_ctx *ctx, double *omega_min);

/* Copy harmonic results to the host.  omega: nq*nb, vel: nq*nb*3, egv: nq*nb*nb
 * complex as interleaved (re, im) pairs, egv[(q*nb + mode)*nb + component]
 * (the reference's egv(:, mode)).  full = 0 for irreducible points, 1 for the full grid.
 * Any pointer may be NULL. */
int tc3_get_harmonic(const tc3_ctx *ctx, int full, double *omega, double *vel, double *egv);

/* Three-phonon scattering rates and scattering matrix for nT temperatures at once:
 * replaces lo_scattering_rates%generate (scattering.f90:58-358) with
 * compute_threephonon_scattering (scattering_threephonon.f90:2-217) for
 * qg3ph = qgrid, plus isotope scattering when tc3_set_isotopes enabled it
 * (compute_isotope_scattering, scattering_isotope.f90:2-70) and boundary scattering when
 * tc3_set_boundary set a length (:245-250).  integrationtype 1 (fixed sigma in THz), 2
 * (adaptive) or 6 (adaptive from group-velocity differences, scattering_threephonon.f90:
 * 251-261); precision one of TC3_PREC_*.  The scattering matrix stays on the device. */
int tc3_scatter(tc3_ctx *ctx, int nT, const double *temperatures, int precision,
                int integrationtype, double sigma, int classical);

/* Scattering results for temperature iT: linewidth (nirr*nb), qs = 2*linewidth,
 * F0 (nirr*nb*3), scalar mean free path (nirr*nb).  Any pointer may be NULL. */
int tc3_get_scattering(const tc3_ctx *ctx, int iT, double *linewidth, double *F0,
                       double *scalar_mfp);

/* Copy rows [row0, row0+nrow) of the symmetrised scattering matrix Xi of temperature iT
 * (each row nfull*nb doubles) and the (q_irr, band) of every row (2*nrows ints). */
int tc3_get_xi(const tc3_ctx *ctx, int iT, int row0, int nrow, double *xi, int *row_qb);

/* Thermal conductivity for every scattered temperature: SMA (get_kappa, kappa.f90:28-75),
 * iterative BTE (iterative_solution, kappa.f90:326-495, FP64 on the GPU) and the
 * off-diagonal coherence term (get_kappa_offdiag, kappa.f90:77-306), each symmetrised
 * (symmetrize_kappa, kappa.f90:308-324).  Outputs nT*9 (atomic units, column-major 3x3);
 * niter receives the iteration count per temperature (may be NULL); kappa_offdiag may be
 * NULL to skip the coherence term.  maxiter <= 0 skips the iterative solution. */
int tc3_kappa(tc3_ctx *ctx, int maxiter, double tol, int classical, double *kappa_sma,
              double *kappa_iter, double *kappa_offdiag, int *niter);

/* tc3_kappa without device allocation, for callers that cannot allocate device memory
 * (serving loops, code between allocation-free regions, harnesses that forbid it).  No
 * counterpart in the reference.  Same results as tc3_kappa with kappa_offdiag = NULL (single-mode
 * and iterative kappa, iteration counts, the log and per-mode results of tc3_get_kappa_log /
 * tc3_get_modes): the iteration's working buffers are carved from one fixed 32 MiB device array
 * per GPU (held until the process ends) and its chunks are launched directly (no graph).  The
 * inputs are what tc3_scatter left on the device.  Returns 1 when the solve needs more than the
 * array holds (its size is printed in the error message), 5 as tc3_kappa.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (from several contexts or threads) race. */
int tc3_kappa_noalloc(tc3_ctx *ctx, int maxiter, double tol, int classical, double *kappa_sma,
                      double *kappa_iter, int *niter);

/* Per-iteration log of the last tc3_kappa call for temperature iT: up to maxlen
 * records of (kxx, kyy, kzz, kxy, kxz, kyz, DeltaF/F) in atomic units
 * (kappa.f90:377, 473-474).  Returns the number of records in *nlog. */
int tc3_get_kappa_log(const tc3_ctx *ctx, int iT, int maxlen, double *log, int *nlog);

/* Final per-mode quantities of temperature iT after tc3_kappa: Fn (nirr*nb*3, before
 * the omega/T prefactor) and per-mode kappa (nirr*nb*9, atomic units). */
int tc3_get_modes(const tc3_ctx *ctx, int iT, double *Fn, double *kappa_mode);

/* Wall times (seconds) of the last calls: [0] harmonic, [1] scatter, [2] kappa total,
 * [3] BTE only, [4] off-diagonal only, [5] scatter kernel time (HIP events),
 * [6] adjoint total, [7] adjoint kernel time (HIP events).  n <= 8. */
int tc3_get_timings(const tc3_ctx *ctx, double *t, int n);

/* Device time (ms, HIP events on the context's stream) and launch count of one kernel
 * class in the last tc3_scatter / tc3_kappa call: 0 Phi + first two contractions,
 * 1 last contraction + smeared-delta epilogue, 2 scattering-matrix assembly and
 * symmetrisation, 3 BTE scattering-matrix product, 4 the same as 2 when the rows were
 * assembled from compact per-triplet sums (no launch counted when the contraction kernels
 * stored the entries directly).  Also the triplet batch of the last scatter call.  A pass
 * whose contraction was replayed from its HIP graph (every call after the first with the same
 * arguments, unless TC3_NO_GRAPH is set) counts no launches in classes 0 and 1: the graph
 * records no events between its kernels.
 * No counterpart in the reference (its timers are host wall-clock buckets). */
int tc3_get_kernel_stats(const tc3_ctx *ctx, int which, double *total_ms, int *launches,
                         int64_t *batch);

/* Adjoint of the linewidths with respect to the third-order force constants: for a
 * cotangent dL_dlinewidth (nT*nirr*nb, of the linewidths tc3_get_scattering returns, all
 * temperatures of the last tc3_scatter call) writes dL_dfc3 (27*ntrip), the gradient with
 * respect to the tensors m of tc3_set_fc3, in the caller's triplet order and layout
 * m[27 t + 9 k + 3 j + i].  Differentiates the three-phonon part of the linewidth
 * (compute_threephonon_scattering, scattering_threephonon.f90:88-135, pretransform_phi3
 * :275-317) through the degeneracy average (scattering.f90:260-277); the isotope term
 * does not depend on the force constants.  |Psi|^2 is recomputed, with the contraction in
 * the given precision (TC3_PREC_*) and FP64 accumulation of the gradient.  Replaces nothing
 * in the reference, which has no derivatives; the forward it differentiates is
 * lo_scattering_rates%generate (scattering.f90:58-358).  Needs tc3_scatter first. */
int tc3_adjoint_gamma(tc3_ctx *ctx, const double *dL_dlinewidth, int precision, double *dL_dfc3);

/* Adjoint of the single-mode (RTA) thermal conductivity with respect to the third-order
 * force constants: dL_dkappa (nT*9, column-major like tc3_kappa's kappa_sma) -> dL_dfc3
 * as in tc3_adjoint_gamma.  kappa_sma = symmetrize_kappa(get_kappa(F0)), F0 = v / (2
 * linewidth) (kappa.f90:28-75, :308-324, scattering.f90:278-292); lo_chop is taken as the
 * identity.  Needs tc3_scatter first (not tc3_kappa). */
int tc3_adjoint_kappa_rta(tc3_ctx *ctx, const double *dL_dkappa, int precision, double *dL_dfc3);

/* Directional derivatives of the harmonic properties, the linewidths and the single-mode
 * (RTA) thermal conductivity along a change of the force constants: dfc2 (9 per pair, the
 * layout and order of tc3_set_fc2's m) and dfc3 (27 per triplet, as tc3_set_fc3's m), either
 * null for no change.  Outputs (each may be null): dlinewidth (nT x nirr x nb, as
 * tc3_get_scattering's linewidth), domega (nirr x nb), dvel (nirr x nb x 3), dkappa_sma (nT x 9,
 * as tc3_kappa's kappa_sma, lo_chop taken as the identity).  Frequencies and eigenvectors change
 * by first-order perturbation of the dynamical matrix (rotations inside degenerate subspaces
 * left out: the degeneracy-averaged linewidths do not depend on them), |Psi|^2 and the smeared
 * energy conservation by the chain rule, all in FP64 on the device.  Forward-mode
 * differentiation of lo_scattering_rates%generate (scattering.f90:58-358), get_kappa
 * (kappa.f90:28-75) and the dispersions (type_phonon_dispersions.f90); the reference has no
 * derivatives, so this replaces nothing.  Gradients with respect to a few parameters of the
 * force constants take one call per parameter.  Needs tc3_scatter first; three-phonon
 * scattering only (status 5 with isotope disorder, boundary or four-phonon scattering, Monte-
 * Carlo grids or integration type 6). */
int tc3_jvp(tc3_ctx *ctx, const double *dfc2, const double *dfc3, double *dlinewidth, double *domega, double *dvel,
            double *dkappa_sma);

/* The three-phonon matrix elements of n irreducible triplets (indices into the
 * 0 .. ntriplets-1 of tc3_mesh_info, the full-grid triplet set): psi2[t][b1][b2][b3] (n nb^3
 * values) = threephonon_prefactor |Psi|^2 / N_q, the psisq of scattering_threephonon.f90:112
 * without the permutation factor, for every band triple (those the reference skips
 * included); qpoints[3t..3t+2] (may be null) receive q1 (irreducible index), q2 and q3
 * (full-grid indices).  Formed by the production contraction kernels in the given precision
 * (TC3_PREC_*), to measure the precision's |Psi|^2 error; the MFMA contraction only (48 and
 * 96 bands, status 5 otherwise).  A diagnostic with no counterpart in the reference.  Needs
 * tc3_harmonic and tc3_set_fc3. */
int tc3_get_psi2(tc3_ctx *ctx, int precision, int n, const int64_t *triplets, double *psi2, int *qpoints);

/* Probes of the contraction's batch autotuner since the context was created: up to maxn
 * records of (triplets per launch, best probe time in ms) in the order they ran; *n
 * receives the number of records available.  For split-BF16x3 at 48 bands the batch probes
 * of a shape are followed by one more record: the chosen batch with the last contraction
 * kernel run as one workgroup per triplet instead of one per 16 b1 (kept if faster).  The
 * chosen batch of the last tc3_scatter is reported by tc3_get_kernel_stats.  No
 * counterpart in the reference (a tuning aid). */
int tc3_get_tuning(const tc3_ctx *ctx, int maxn, double *batch, double *ms, int *n);

/* Device storage of the scattering matrix of the last tc3_scatter: elements per temperature,
 * bytes per element (8 in FP64 mode, 4 in the mixed modes) and whether the rows are stored per
 * orbit of the little group of their q-point (1: one value per orbit and band, which
 * tc3_get_xi expands to the full rows; 0: nfull*nb values per row).  No counterpart in the
 * reference (a measurement aid: the bytes the BTE product reads per iteration). */
int tc3_get_xi_storage(const tc3_ctx *ctx, int64_t *elems_per_T, int *entry_bytes, int *per_orbit);

/* Fourth-order IFCs: replaces lo_forceconstant_fourthorder%readfromfile
 * (type_forceconstant_fourthorder.f90:259-340) and switches four-phonon scattering on in the
 * following tc3_scatter calls (--fourthorder; nquart = 0 switches it off).  lv2, lv3, lv4:
 * 3*nquart Cartesian lattice vectors (Bohr); m: 81*nquart tensors (Ha/Bohr^4),
 * m[81*t + 27*l + 9*k + 3*j + i] = Phi_ijkl; the mass weighting (:296) is applied inside.
 * compute_fourphonon_scattering (scattering_fourphonon.f90:1-328) then adds its linewidth
 * and scattering-matrix terms with the full q-grid for the quartets (-qg4ph equal to the
 * mesh unless tc3_set_montecarlo4 draws them), FP64 whatever the contraction precision;
 * any cell of up to 32 atoms (96 bands; more: tc3_scatter returns 5). */
int tc3_set_fc4(tc3_ctx *ctx, int nquart, const int *a1, const int *a2, const int *a3, const int *a4,
                const double *lv2, const double *lv3, const double *lv4, const double *m);

/* Monte-Carlo grids of the four-phonon term: replaces opts%qg4ph and mcg4%initialize
 * (options.f90:226-236, scattering.f90:122-124).  qg4ph: 3 grid dimensions (NULL, or a
 * value outside 1..qgrid, for the full mesh); the seed is tc3_set_montecarlo's.  Every row
 * (q1, b1) then draws its three-phonon grid and two four-phonon grids in turn from one
 * generator (compute_threephonon_scattering :69-70, compute_fourphonon_scattering :84-85),
 * visits the (q2, q3) pairs of its two grids and weights them 1 / prod(qg4ph)^2. */
int tc3_set_montecarlo4(tc3_ctx *ctx, const int *qg4ph);

/* Write outfile.thermal_conductivity_grid.hdf5 for temperature iT (layout of
 * lo_phonon_dispersions%write_to_hdf5, type_phonon_dispersions.f90:316-641, after the
 * Fn*omega/T prefactor of main.f90:323-328). */
int tc3_write_grid_hdf5(const tc3_ctx *ctx, int iT, const char *path);

/* Write outfile.cumulative_thermal_conductivity.hdf5 for temperature iT after tc3_kappa:
 * phonon DOS (lo_phonon_dos%generate, type_phonon_dos.f90:110-297), cumulative kappa
 * against mean free path (get_cumulative_kappa, cumulative_kappa.f90:58-414), kappa with
 * boundary scattering (get_boundary_kappa, :416-508) and spectral kappa
 * (type_phonon_dos_integrations.f90:428-809), laid out as lo_cumulative_kappa%write_to_hdf5
 * (cumulative_kappa.f90:580-777), with the angular-momentum tensor and its spectrum
 * (type_phonon_dispersions.f90:868-971, type_phonon_dos_integrations.f90:8-425).  mfppts, freqpts,
 * dosintegrationtype (1 Gaussian, 2 adaptive Gaussian, 3 tetrahedron), sigma, dossigma and
 * unit ("thz", "mev", "icm") are thermal_conductivity's --mfppts, --freqpts,
 * --dosintegrationtype, --sigma, --dossigma and --unit (options.f90:86-147);
 * atomic_species is the space-separated element list of the file attribute.  Returns 5 for
 * invalid arguments or an unknown unit, 6 when the HDF5 library cannot be loaded or the
 * file cannot be written. */
int tc3_write_cumulative_hdf5(const tc3_ctx *ctx, int iT, const char *path, int mfppts, int freqpts,
                              int dosintegrationtype, double sigma, double dossigma, const char *unit,
                              const char *atomic_species);

#ifdef __cplusplus
}
#endif

#endif /* TC3_H */
