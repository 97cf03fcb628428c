/* fdtd_hip.h — C ABI of the MI355X FDTD hot path.
 *
 * One opaque grid handle owns all device state of one simulation: SoA field
 * arrays, CPML auxiliary (psi) slabs, per-axis CPML profiles, a uint8
 * material-id volume with a material table, a dense dispersive box (D, P, J),
 * the TFSF incident line, running-DFT accumulators and probe traces.
 * fdtd_step() runs the reference's whole step composition, which the
 * reference spells as one functor sweep per stage (tests/testTM.cpp:102-113,
 * tests/test3D.cpp:146-171):
 *   D curl -> soft sources -> CPML(E) -> TFSF updateD + aux step -> E material
 *   -> B curl -> sources -> CPML(H) -> TFSF updateB -> H material
 * plus the running DFT and probe capture.  All pointers here are host
 * pointers; the library copies to/from device memory.  Every function returns
 * 0 on success or a negative FDTD_E* code; fdtd_last_error() then holds a
 * thread-local message.  No exceptions cross this boundary.
 *
 * Index convention: cell (i,j,k), x fastest: q = i + nx*(j + ny*k).  The
 * outermost layer of cells (index 0 and N-1 on every axis of the mode) is
 * never updated, as in the reference (tests/testTM.cpp:103).
 */
#ifndef FDTD_HIP_H
#define FDTD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum fdtd_status {
  FDTD_OK = 0,
  FDTD_EINVAL = -1,   /* bad argument */
  FDTD_ENOMEM = -2,   /* device allocation failed */
  FDTD_EDEVICE = -3,  /* HIP runtime error */
  FDTD_ESTATE = -4    /* call not valid in the grid's current state */
};

/* Modes, as in include/FDTDConstants.hpp:424-445: TEM (Ez, Hy along x; a
 * line of nx cells, ny = nz = 1), TM (Ez, Hx, Hy), TE (Ex, Ey, Hz), ThreeD. */
enum fdtd_mode { FDTD_TEM = 1, FDTD_TM = 2, FDTD_3D = 3, FDTD_TE = 4 };
enum fdtd_precision { FDTD_F32 = 0, FDTD_F64 = 1 };
/* field codes (M, K: magnetization and magnetic current of magnetic poles) */
enum fdtd_field { FDTD_D = 0, FDTD_E = 1, FDTD_B = 2, FDTD_H = 3, FDTD_P = 4, FDTD_J = 5, FDTD_M = 6, FDTD_K = 7 };
/* OR'ed into the field of fdtd_upload_field / fdtd_download_field: the
 * imaginary part of E or H of a complex grid (one with a Bloch-periodic face) */
#define FDTD_IMAG 16
/* material pole kinds */
enum fdtd_pole { FDTD_DRUDE = 1, FDTD_LORENTZ = 2, FDTD_CONDUCTIVE = 3 };
/* boundary conditions: the kinds of fdtd::Boundary (include/BoundaryUpdates.hpp:17-26) */
enum fdtd_boundary {
  FDTD_BC_NONE = 0,          /* ghost layer left alone (stays as uploaded; zero by default) */
  FDTD_BC_PEC = 1,
  FDTD_BC_PMC = 2,
  FDTD_BC_SYMMETRIC = 3,
  FDTD_BC_ANTISYMMETRIC = 4,
  FDTD_BC_PERIODIC = 5,
  FDTD_BC_BLOCH = 6          /* set by fdtd_set_bloch_boundary (a complex phase factor) */
};

#define FDTD_MAX_MATERIALS 256
#define FDTD_MAX_POLES 3
#define FDTD_MAX_SOURCES 16
#define FDTD_MAX_DFT 8
#define FDTD_MAX_PROBES 32
#define FDTD_MAX_PLANES 6
/* the side of a face: Orientation::MIN / MAX (include/FDTDConstants.hpp) */
enum fdtd_side { FDTD_MIN = 0, FDTD_MAX = 1 };

typedef struct fdtd_grid fdtd_grid;

typedef struct {
  int mode;      /* FDTD_TEM, FDTD_TM, FDTD_TE or FDTD_3D */
  int precision; /* FDTD_F32 or FDTD_F64: storage and arithmetic type */
  int nx, ny, nz; /* cells per axis including the outer layer; nz = 1 for TM and TE, ny = nz = 1 for TEM */
  double dx;     /* cell size [m] */
  double dt;     /* time step [s] */
} fdtd_grid_desc;

/* Replaces the user-built std::vector<YeeCell> grid + neighbour pointers of
 * tests/test3D.cpp:33-49 (YeeCell.hpp:74-97, YeeFields.hpp:11-174). */
int fdtd_grid_create(const fdtd_grid_desc* desc, fdtd_grid** out);
int fdtd_grid_destroy(fdtd_grid* g);
/* hipStream_t to launch on (NULL = the default stream). */
int fdtd_set_stream(fdtd_grid* g, void* hip_stream);

/* NewPML::setPMLParameters (include/PML.hpp:84-102) for every index of one
 * axis: K, S, A are arrays of the axis length; ft 0 = electric, 1 = magnetic.
 * The reference stores these per cell; here they are per-axis profiles. */
int fdtd_set_pml_axis(fdtd_grid* g, int axis, int ft, const double* K, const double* S, const double* A);
/* PMLParameterModel (include/PML.hpp:368-393) with the reference's placement
 * idiom (tests/testTM.cpp:46-71): n graded cells at both ends of `axis`. */
int fdtd_set_pml_idiom(fdtd_grid* g, int axis, int n, double m, double ma, double smax, double kmax, double amax);

/* Material table entry `id`: replaces MaterialPair/MaterialMap
 * (include/MaterialMap.hpp:45-284) and the Vacuum/Constant/Drude/Lorentz
 * update functors (include/DispersiveMaterials.hpp:342-497, 692-884,
 * 898-1217).  poles: npoles x {kind, freq [rad/s], gamma [1/s], weight}.
 * One Drude pole of weight 1 is the reference's DrudeUpdate; one Lorentz pole
 * of weight 1 is its LorentzUpdate (which never reads its delta); one
 * FDTD_CONDUCTIVE pole {3, sigma [S/m], unused, unused} is its
 * ConductiveUpdate (508-680): P = (P + f D)/(1 + f), f = dt sigma/(eps_r eps0),
 * E = (D - P)/(eps_r eps0). */
int fdtd_set_material(fdtd_grid* g, int id, double eps_r, double mu_r, int npoles, const double* poles);
/* The magnetic side of material `id`: replaces the reference's DrudeUpdate /
 * LorentzUpdate / ConductiveUpdate on FieldType::Magnetic (the same
 * recursions on B, H, M, K with mu0 and mu_r for eps0 and eps_r: the *Call
 * structs of include/DispersiveMaterials.hpp, val = mu0; M and K are the
 * DoubleMagnetization members, 86-110).  poles as for fdtd_set_material
 * (FDTD_CONDUCTIVE: freq = the magnetic conductivity); mu_r is
 * fdtd_set_material's.  npoles = 0 clears them.  Grids holding such cells run
 * the general two-pass instances; batches, the adjoint gradient, complex
 * (Bloch) grids and self-magnetized Drude materials refuse them. */
int fdtd_set_material_magnetic(fdtd_grid* g, int id, int npoles, const double* poles);
/* Material `id` as a magnetized Drude plasma: replaces MagnetizedDrudeUpdate
 * (bias != NULL: the fixed field bias[3] in tesla, MagnetizedDrudeStorage,
 * include/DispersiveMaterials.hpp:1619-1726) and SelfMagnetizedDrudeUpdate
 * (bias == NULL: the cell's own magnetic flux density B = mu_r mu0 H,
 * SelfMagnetizedDrudeStorage, 1898-2006).  In 3-D the three components of
 * P and J advance together by MagnetizedDrudeUpdate3D::implicit_update
 * (1283-1340); in TM the reference falls back to DrudeUpdate per component
 * (MagnetizedDrudeCall, 1508-1537), and so does this material.  K = eps_r,
 * Freq = omega_p [rad/s], Gamma [1/s]. */
int fdtd_set_magnetized_drude(fdtd_grid* g, int id, double eps_r, double mu_r, double freq, double gamma,
                              const double* bias);
/* Material `id` as a fluid Drude plasma whose plasma frequency and collision
 * rate are those of each cell: omega_p = sqrt(q_e^2/(m_e eps0)) sqrt(ne),
 * gamma = nu_m, K = 1 (FluidDrudeStorage, include/DispersiveMaterials.hpp:
 * 1160-1199).  magnetization 0: FluidDrudeUpdate (the Drude update per
 * component); 1: MagnetizedFluidDrudeUpdate in the fixed field bias[3] [T]
 * (1728-1814); 2: SelfMagnetizedFluidDrudeUpdate in the cell's own B
 * (1816-1893).  In TM all three are the Drude update per component, as in the
 * reference.  The per-cell values come from fdtd_set_fluid_state. */
int fdtd_set_fluid_drude(fdtd_grid* g, int id, double mu_r, int magnetization, const double* bias);
/* The cells' electron density ne [1/m^3] and momentum-transfer collision rate
 * nu_m [1/s] (the ne() / nu_m() accessors the reference's fluid storages read
 * from a cell), nx*ny*nz each, read where the material is a fluid Drude one.
 * May be called again between steps (a coupled fluid model). */
int fdtd_set_fluid_state(fdtd_grid* g, const double* ne, const double* nu_m);
/* per-cell material id, nx*ny*nz bytes */
int fdtd_set_matid(fdtd_grid* g, const uint8_t* ids);

/* make_soft_source + SoftSource::apply (include/Source.hpp:183-261): at step
 * n, table[n] is added to D_comp (field FDTD_D) or B_comp (FDTD_B) of cell
 * (i,j,k) right after the curl, before CPML and the material step (a lasting
 * source: testTM.cpp:105), or to E_comp (FDTD_E) or H_comp (FDTD_H) right
 * after the material step, before the boundaries (the placement of Aux1D's
 * own source, Source.hpp:421-425: the next material step recomputes E from
 * D, H from B, so the value acts on one curl of the other field only).
 * E and H are stored without D and B, so a value of E or H uploaded between
 * steps stands for the field including the last step's source value, which
 * the next step takes back off.  Steps n >= ntable add nothing. */
int fdtd_add_source(fdtd_grid* g, int field, int comp, int i, int j, int k, const double* table, int ntable);

/* TFSF box = make_tfsf_plane x 2*dim faces sharing one Aux1D history
 * (include/Source.hpp:280-634).  Faces lie on cell layers box[0],box[1]
 * (x), box[2],box[3] (y), box[4],box[5] (z).  kvec is normalised as in
 * TFSFPlane's constructor.  pol multiplies the aux E (updateB), polH the aux
 * H (updateD).  origin: cell-unit coordinates of the location-functor origin
 * (cell centre r = (i - origin_x) dx, ...).  aux_ncells sizes the Aux1D line;
 * signal[n] is its soft-source value at its n-th step. */
int fdtd_set_tfsf(fdtd_grid* g, const int box[6], const double kvec[3], const double pol[3], const double polH[3],
                  const double origin[3], int aux_ncells, const double* signal, int nsignal);

/* One TFSF plane = make_tfsf_plane<Mode, Dir, Orient>(begin, end, loc, k, aux)
 * + TFSFPlane::updateD / updateB (include/Source.hpp:472-634) with its own
 * Aux1D: normal along `axis` (0 x, 1 y, 2 z) pointing to `side` (FDTD_MIN:
 * -1, FDTD_MAX: +1); the plane's cells are the box [x0,x1)x[y0,y1)x[z0,z1)
 * of `cells`, one layer thick along the axis (the iterator range; TM:
 * z range ignored).  Every cell of it gets, in each step, the correction of
 * every flux but the normal one, on D after the E-side CPML and on B after
 * the H-side CPML, in the order the planes were added (after the box's).  The
 * location functor is r = (cell - origin) dx.  kvec is normalised; pol / polH
 * as for fdtd_set_tfsf (the reference's defaults per mode,
 * Source.hpp:573-584, are the caller's to pass).  At most FDTD_MAX_PLANES. */
int fdtd_add_tfsf_plane(fdtd_grid* g, int axis, int side, const int cells[6], const double kvec[3],
                        const double pol[3], const double polH[3], const double origin[3], int aux_ncells,
                        const double* signal, int nsignal);

/* Rewrite or extend a per-step table between steps, for a time loop that
 * computes each step's value as it goes (SoftSource::apply(t) in the loop,
 * include/Source.hpp:225-229; the signal Aux1D::time_step samples, 414-425):
 * values[q] becomes the entry of step step0 + q of
 *   FDTD_SIGNAL_SOURCE: the table of source `index` (sources are numbered
 *                       0, 1, ... in the order of fdtd_add_source calls),
 *   FDTD_SIGNAL_TFSF:   the TFSF box's signal (index 0),
 *   FDTD_SIGNAL_PLANE:  the signal of plane `index` (order of
 *                       fdtd_add_tfsf_plane calls).
 * The table grows as needed; entries never written are 0.  Waits for the
 * grid's queued steps first. */
enum fdtd_signal { FDTD_SIGNAL_SOURCE = 0, FDTD_SIGNAL_TFSF = 1, FDTD_SIGNAL_PLANE = 2 };
int fdtd_set_signal(fdtd_grid* g, int kind, int index, long step0, const double* values, int n);

/* Running DFT over the half-open region [i0,i1)x[j0,j1)x[k0,k1) (no
 * counterpart in the reference: include/DataCapture.hpp:61-191 is empty).
 * comp_mask bit c selects E_c (c<3) or H_{c-3}.  phaseE/phaseH:
 * [nphase][nfreq][2] = (cos, -sin)(w t) dt with E sampled after step n at
 * t=(n+1)dt and H at t=(n+1.5)dt.  *idx receives the monitor index. */
int fdtd_add_dft(fdtd_grid* g, const int region[6], int comp_mask, int nfreq, const double* phaseE,
                 const double* phaseH, int nphase, int* idx);
/* [ncomp][nfreq][2][cells of region] as float64 */
int fdtd_get_dft(fdtd_grid* g, int idx, double* out);

/* Probe: field value at one cell recorded after every step. */
int fdtd_add_probe(fdtd_grid* g, int field, int comp, int i, int j, int k);
/* out: [nsteps][nprobes] float64, nsteps <= max_steps; *nsteps_out set */
int fdtd_get_probes(fdtd_grid* g, double* out, int max_steps, int* nsteps_out);
/* the imaginary part of the probes of a complex grid (same layout) */
int fdtd_get_probes_imag(fdtd_grid* g, double* out, int max_steps, int* nsteps_out);

/* Host<->device copies of one component (nx*ny*nz float64 values).  D and B
 * read back as eps_r eps0 E and mu_r mu0 H outside the cells of pole
 * materials; P, J (electric poles) and M, K (magnetic poles) take `pole`. */
int fdtd_upload_field(fdtd_grid* g, int field, int comp, int pole, const double* host);
int fdtd_download_field(fdtd_grid* g, int field, int comp, int pole, double* host);

/* Advance nsteps whole time steps (asynchronous on the grid's stream). */
int fdtd_step(fdtd_grid* g, int nsteps);
int fdtd_synchronize(fdtd_grid* g);
/* steps taken so far */
long fdtd_step_count(const fdtd_grid* g);

/* Schedule and tile of the per-step kernels.  bx >= 2: two passes (E, then
 * H) with bx*by threads per block (a multiple of 64, at most 256) covering
 * bx x by cells of an xy plane, each thread marching kc z planes.  bx = 0: one
 * fused pass per step, workgroups of 254 columns x `by` rows (1..3) sharing
 * neighbours in LDS.  bx = 1: one fused pass per step, each wavefront owning
 * 62 columns x `by` rows (2, 4 or 6) and exchanging neighbours by lane
 * shuffles.  Fused passes keep two copies of the stepped state.  Results do
 * not depend on the schedule or tile; only speed does. */
int fdtd_set_tile(fdtd_grid* g, int bx, int by, int kc);
int fdtd_get_tile(fdtd_grid* g, int* bx, int* by, int* kc);
/* Times each candidate tile over trial_steps steps, keeps the fastest and
 * restores the complete device state and step count afterwards (it
 * snapshots every allocation of the grid, so it needs as much free device
 * memory again).  The chosen tile is returned through bx, by, kc. */
int fdtd_autotune(fdtd_grid* g, int trial_steps, int* bx, int* by, int* kc);

/* Surface charge density and surface currents of the current state on the
 * cells of region [x0,x1)x[y0,y1)x[z0,z1) (interior cells; TM and TE: z
 * range [0,1)): replaces UpdateSurfaceChargeDensity / UpdateSurfaceCurrentE /
 * UpdateSurfaceCurrentH (include/SurfaceQuantities.hpp:169-356) applied to
 * those cells.  out: quantity-major, x fastest; 3-D 7 quantities rho, Jex,
 * Jey, Jez, Jhx, Jhy, Jhz; TM 3: Jez, Jhx, Jhy (the reference's TM charge
 * density functor does nothing); TE 4: rho, Jex, Jey, Jhz.  FP64 arithmetic; synchronises.  On a
 * complex grid (Bloch face) these are the quantities of the real part. */
int fdtd_surface_quantities(fdtd_grid* g, const int region[6], double* out);

/* The same quantities as fdtd_surface_quantities, without allocating: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving or capture loops, harnesses that forbid hipMalloc).
 * The device staging is a fixed array of the library on each GPU
 * (7 x FDTD_SURFACE_STATIC_MAX_CELLS doubles, 3.5 MiB, held until the process
 * ends); regions of more cells are refused (FDTD_EINVAL), and so is a grid
 * whose device set-up has not run yet (FDTD_ESTATE: step it, or call
 * fdtd_surface_quantities, once first).  Calls on one GPU share that array:
 * calls that overlap in time on one GPU race; run them one after another. */
#define FDTD_SURFACE_STATIC_MAX_CELLS 65536
int fdtd_surface_quantities_static(fdtd_grid* g, const int region[6], double* out);

/* Boundary condition on one face (axis 0..ndim-1, side 0 = min, 1 = max).
 * Replaces make_pec_boundary / make_pmc_boundary / make_symmetric_boundary /
 * make_antisymmetric_boundary / make_periodic_boundary + BoundaryData::
 * updateE / updateH (include/BoundaryUpdates.hpp:527-832): after each E (H)
 * update, every E (H) component of the face's ghost layer (index 0 or L-1)
 * becomes the reference's mask (+-1 per component) times the same component
 * of the adjacent interior layer (1 or L-2), or, periodic, of the layer one
 * period away (L-2 for ghost 0, 1 for ghost L-1).  Tangential range: the
 * interior cells.  Set before the first step; not with batches or the
 * adjoint gradient (FDTD_EINVAL there). */
int fdtd_set_boundary(fdtd_grid* g, int axis, int side, int kind);
/* A Bloch-periodic face: make_bloch_periodic_boundary (include/BoundaryUpdates.hpp:
 * 845-885), which needs complex fields.  The grid's fields become complex
 * (the excitation, sources and TFSF, stays real): the ghost layer of the face
 * takes factor * (the periodic image) for E and conj(factor) * (image) for H,
 * the reference's decorator (mask 1 for every component).  The imaginary
 * part is read with FDTD_IMAG and fdtd_get_probes_imag; fdtd_get_dft returns
 * the spectrum of the complex field.  Before stepping; complex grids run the
 * two-pass schedule and cannot be batched or differentiated.  The two parts
 * are stepped separately, which needs every update linear with real
 * coefficients: a 3-D grid with a self-magnetized (Drude or fluid) material
 * is refused. */
int fdtd_set_bloch_boundary(fdtd_grid* g, int axis, int side, double factor_re, double factor_im);

/* Capture each chunk of steps (up to 256) in a hipGraph once and replay it
 * for later chunks of the same shape (off by default: plain launches).  Pays
 * off where launches dominate (small grids).  A grid on the default stream
 * gets a stream of its own for capture.  Results are unchanged. */
int fdtd_set_graph(fdtd_grid* g, int on);

/* Per-kernel timing with hipEvents on the grid's stream (off by default).
 * fdtd_kernel_time returns accumulated milliseconds and launches for
 * kernel id 0 = E pass, 1 = H pass, 2 = aux line, 3 = fused pass, 4 = DFT,
 * 5 = adjoint E pass, 6 = adjoint H pass, 7 = Bloch ghost layers. */
int fdtd_set_profiling(fdtd_grid* g, int on);
int fdtd_kernel_time(fdtd_grid* g, int kernel, double* ms, long* launches);

/* Design box [i0,i1)x[j0,j1)x[k0,k1) of interior cells whose relative
 * permittivity is given per cell (eps_r, x fastest), replacing the material
 * table's eps_r there; such cells are non-dispersive, mu_r stays the
 * material's.  The reference sets eps_r per material only
 * (include/DispersiveMaterials.hpp:342-359); this per-cell form exists for
 * the adjoint gradient below.  Calling it again replaces the values. */
int fdtd_set_design(fdtd_grid* g, const int box[6], const double* eps_r);

/* Discrete adjoint gradient (no counterpart in the reference).  From the
 * initial state (step 0) steps the grid nsteps forward, then returns the
 * objective J = sum over the cells of DFT monitor `dft`, the E components of
 * comp_mask (bit c = E_c, recorded by that monitor) and its frequencies of
 * |F|^2, and grad[cell] = dJ/d eps_r for every design cell (same order as
 * fdtd_set_design).  Reverse steps are the transpose of the forward step
 * (curl, CPML psi, per-cell material including the P / J recursions of
 * Drude, Lorentz and conductive poles; sources and TFSF terms are
 * state-independent).  The design cells' E increments are kept for segments
 * of checkpoint_every steps: the last segments that fit history_mb megabytes
 * (at least one) are recorded during the forward pass, every earlier segment
 * restarts from a checkpoint of E, H, psi and the pole cells' D, P, J taken
 * at its first step and is
 * recomputed just before its reverse steps (each step over the design box
 * grown by one cell per step left in the segment, on grids without
 * boundary ghost writes or complex fields).  On return the grid holds the
 * forward state after nsteps steps and the monitor holds F.  3-D grids;
 * magnetized and fluid Drude materials are refused (FDTD_EINVAL); design
 * cells are non-dispersive (their own eps_r). */
int fdtd_adjoint_eps_gradient(fdtd_grid* g, int nsteps, int checkpoint_every, double history_mb, int dft,
                              int comp_mask, double* objective, double* grad);
/* Tile of the gradient's reverse sweep and of its recomputed segments (the
 * forward pass keeps fdtd_set_tile's): the two run side by side on two
 * streams, where other shapes share the GPU best.  (0, 0, 0) = the grid's
 * tile.  No counterpart in the reference. */
int fdtd_set_adjoint_tile(fdtd_grid* g, int bx, int by, int kc);
/* The same gradient for n independent grids (designs) at once: each runs on
 * its own host thread and stream (a non-blocking one is made for a grid on
 * the default stream), so the designs share the GPU.  objectives[q] and
 * grads[q] receive grid q's results; per design they equal single calls.
 * No counterpart in the reference. */
int fdtd_adjoint_eps_gradient_batch(fdtd_grid* const* grids, int n, int nsteps, int checkpoint_every,
                                    double history_mb, int dft, int comp_mask, double* objectives,
                                    double* const* grads);

/* Batches: independent 3-D grids of one precision stepped together, one
 * launch per pass covering the tiles of every grid, each chunk of steps
 * captured once in a hipGraph and replayed.  The grids stay owned by the
 * caller and must outlive the batch; while batched they run on the batch's
 * stream and must not be stepped on their own.  No counterpart in the
 * reference (its user loop steps one grid, tests/test3D.cpp:146-171). */
typedef struct fdtd_batch fdtd_batch;
int fdtd_batch_create(fdtd_grid* const* grids, int n, fdtd_batch** out);
int fdtd_batch_destroy(fdtd_batch* b);
int fdtd_batch_set_graph(fdtd_batch* b, int on);
/* Two-pass tile of the batch's passes (as fdtd_set_tile with bx >= 2; the
 * default is 64 x 4 x 2). */
int fdtd_batch_set_tile(fdtd_batch* b, int bx, int by, int kc);
/* Persistent passes: wg_per_cu workgroups per compute unit (rounded to a
 * multiple of 8, at most one per tile) walk the tile table in every pass,
 * each staging a grid's material table once per run of that grid's tiles;
 * 0 (the default) launches one workgroup per tile.  Bitwise the same
 * results either way.  No counterpart in the reference. */
int fdtd_batch_set_persistent(fdtd_batch* b, int wg_per_cu);
/* Times candidate tiles over trial_steps steps of the whole batch, then the
 * best tile's persistent forms (2, 4, 6, 8 workgroups per compute unit),
 * keeps the fastest and restores every grid's device state and step count. */
int fdtd_batch_autotune(fdtd_batch* b, int trial_steps, int* bx, int* by, int* kc);
int fdtd_batch_step(fdtd_batch* b, int nsteps);
int fdtd_batch_synchronize(fdtd_batch* b);

const char* fdtd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
