/* refcpu — FP64 CPU restatement of the reference's per-time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  Only tests/, the smoke check and the
 * CPU-baseline leg of bench.py load this library; the product path never does.
 *
 * It restates, cell by cell and in the reference's own operation order:
 *   YeeUpdate curl           include/YeeUpdates.hpp:343-405
 *   UpdatePML (psi + flux)   include/PML.hpp:296-355
 *   NewPML coefficients      include/PML.hpp:84-102
 *   PMLParameterModel        include/PML.hpp:368-393
 *   Constant/Vacuum update   include/DispersiveMaterials.hpp:354-359
 *   Drude implicit update    include/DispersiveMaterials.hpp:979-998
 *   Lorentz implicit update  include/DispersiveMaterials.hpp:710-735
 *   Aux1D                    include/Source.hpp:280-458
 *   TFSFPlane updateD/B      include/Source.hpp:489-619
 *   step composition         tests/testTM.cpp:102-113, tests/test3D.cpp:146-171
 * plus the build-defined pieces the reference lacks (running DFT, multi-pole
 * dispersion with a Lorentz weight, per-component TFSF face ranges).
 */
#ifndef FDTD_REFCPU_H
#define FDTD_REFCPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RC_MAXPOLES 3
#define RC_MAXMAT 256
#define RC_MAXPLANE 6

/* modes: which components exist (FDTDConstants.hpp:424-445, 844-874) */
enum { RC_TEM = 1, RC_TM = 2, RC_3D = 3, RC_TE = 4 };

typedef struct rc_sim rc_sim;
typedef struct rc_aux rc_aux;

/* constants of include/FDTDConstants.hpp:394-398 */
double rc_eps0(void);
double rc_mu0(void);
double rc_c0(void);
double rc_imp0(void);

/* NewPML::setPMLParameters closed form: out = {B, C, F, G} */
void rc_pml_coeffs(double K, double S, double A, double dt, double out[4]);
/* PMLParameterModel::K/S/A at x: out = {K, S, A} */
void rc_pml_model(double x, double m, double ma, double smax, double kmax, double amax, double out[3]);
/* placement idiom along one axis of length L: outE/outH are [L][7] = K,S,A,B,C,F,G
 * (defaults K=1,S=A=B=C=G=0,F=1 where not placed) */
void rc_pml_profile(int L, int n, double m, double ma, double smax, double kmax, double amax, double dt,
                    double* outE, double* outH);

/* Aux1D<double> (TEM line, 15-cell CPML each end, source cell 20) */
rc_aux* rc_aux_create(int ncells, double dx, double dt);
void rc_aux_destroy(rc_aux* a);
int rc_aux_length(const rc_aux* a);
void rc_aux_step(rc_aux* a, double source_value); /* one Aux1D::time_step */
double rc_aux_interp_E(const rc_aux* a, double loc);
double rc_aux_interp_H(const rc_aux* a, double loc);
void rc_aux_state(const rc_aux* a, double* Ez, double* Hy); /* L values each */

/* grid simulation */
rc_sim* rc_create(int mode, int nx, int ny, int nz, double dx, double dt);
void rc_destroy(rc_sim* s);
/* per-axis, per-index PML coefficients; coeffs is [L][4] = B,C,F,G; ft 0=E 1=H */
int rc_set_pml_axis(rc_sim* s, int axis, int ft, const double* coeffs);
/* material table: eps_r, mu_r, npoles, then per pole {kind(1 Drude,2 Lorentz), freq, gamma, weight} */
int rc_set_material(rc_sim* s, int id, double eps_r, double mu_r, int npoles, const double* poles);
/* magnetic side of material id: poles (kind, freq, gamma, weight) on B, H, M, K; rc_get_field field 8 = M, 9 = K */
int rc_set_material_magnetic(rc_sim* s, int id, int npoles, const double* poles);
/* magnetized Drude material (bias NULL: self-magnetized by the cell's own B);
 * in 3-D the coupled update, otherwise the Drude pole {1, freq, gamma, 1} */
int rc_set_magnetized_drude(rc_sim* s, int id, double eps_r, double mu_r, double freq, double gamma,
                            const double* bias);
/* fluid Drude material (K = 1; magnetization 0 none, 1 bias, 2 the cell's own
 * B) and the per-cell electron density / collision rate it reads */
int rc_set_fluid_drude(rc_sim* s, int id, double mu_r, int magnetization, const double* bias);
int rc_set_fluid_state(rc_sim* s, const double* ne, const double* nu_m);
int rc_set_matid(rc_sim* s, const uint8_t* ids); /* nx*ny*nz, x fastest */
/* soft source: adds table[n] to D_comp (field 0) or B_comp (field 1) at cell after the curl of step n,
 * or to E_comp (field 2) / H_comp (field 3) after the material step of step n */
int rc_add_source(rc_sim* s, int field, int comp, int i, int j, int k, const double* table, int ntable);
/* TFSF box; faces at x0,x1 (etc); kvec/pol/polH 3-vectors; origin (cell units); aux line of ncells;
 * signal[n] is the aux source value at its n-th step */
int rc_set_tfsf(rc_sim* s, const int box[6], const double kvec[3], const double pol[3], const double polH[3],
                const double origin[3], int aux_ncells, const double* signal, int nsignal);
/* single TFSF plane with its own aux line: normal along axis (side 0 MIN: -1, 1 MAX: +1), cells
 * [x0,x1)x[y0,y1)x[z0,z1); every flux but the normal one of every cell is corrected */
int rc_add_tfsf_plane(rc_sim* s, int axis, int side, const int cells[6], const double kvec[3], const double pol[3],
                      const double polH[3], const double origin[3], int aux_ncells, const double* signal,
                      int nsignal);
/* running DFT over a box region; comps bitmask (bit c: E_c for c<3, H_{c-3}); phases [nsteps][nfreq][2]
 * for E and for H (already multiplied by dt) */
int rc_add_dft(rc_sim* s, const int region[6], int comp_mask, int nfreq, const double* phaseE,
               const double* phaseH, int nphase);
int rc_get_dft(const rc_sim* s, int idx, double* out); /* [ncomp][nfreq][2][ncells] */
int rc_add_probe(rc_sim* s, int field, int comp, int i, int j, int k);
int rc_get_probes(const rc_sim* s, double* out, int max_steps); /* [steps][nprobes] */

/* field: 0 D, 1 E, 2 B, 3 H, 4 P(pole p), 5 J(pole p) ; comp 0..2; 6 psi E, 7 psi H with comp = 3 I + J */
int rc_get_field(const rc_sim* s, int field, int comp, int pole, double* out);
int rc_set_field(rc_sim* s, int field, int comp, int pole, const double* in);
int rc_step(rc_sim* s, int nsteps);

/* ---- build-defined adjoint (the reference has none) */
/* per-cell eps_r over box [i0,i1)x[j0,j1)x[k0,k1) (x fastest); design cells are non-dispersive */
int rc_set_design(rc_sim* s, const int box[6], const double* eps_r);

/* boundary conditions (BoundaryUpdates.hpp:17-26 names; numbering the build's) */
enum { RC_BC_NONE = 0, RC_BC_PEC = 1, RC_BC_PMC = 2, RC_BC_SYMMETRIC = 3, RC_BC_ANTISYMMETRIC = 4, RC_BC_PERIODIC = 5,
       RC_BC_BLOCH = 6 };
int rc_set_boundary(rc_sim* s, int axis, int side, int type);
/* Bloch-periodic face with phase factor fr + i fi: the fields become complex
 * (field code | 16 reads the imaginary part; rc_get_probes_imag) */
int rc_set_bloch_boundary(rc_sim* s, int axis, int side, double fr, double fi);
int rc_get_probes_imag(const rc_sim* s, double* out, int max_steps);
double rc_bc_mask(int type, int ft, int comp, int axis);
/* the ghost-layer writes of one field type (0: E, 1: H), as done inside rc_step */
void rc_apply_boundaries(rc_sim* s, int ft);

/* surface charge density and surface currents on a region (SurfaceQuantities.hpp) */
int rc_surface_quantities(const rc_sim* s, const int region[6], double* out);
/* E = D/(eps_r eps0), H = B/(mu_r mu0) on interior cells (non-dispersive) */
void rc_sync_derived(rc_sim* s);
/* Gradient from step 0 over nsteps of J = sum over the cells of DFT monitor `dft`, its E components in
 * comp_mask and its frequencies of |F|^2, with respect to every design cell's eps_r. */
int rc_adjoint_eps_gradient(rc_sim* s, int nsteps, int dft, int comp_mask, double* objective, double* grad);
/* Transposed step, for the dot-product test: adjoint state D_bar (field 0), B_bar (2), psi_bar (6 E, 7 H;
 * comp = 3 I + J), pole P_bar / J_bar (8 / 9; comp = 3 p + I).  reset zeroes it; step applies the transpose of the homogeneous forward step n;
 * finish folds the adjoint of the step's input H = B/(mu_r mu0) into B_bar. */
void rc_adjoint_reset(rc_sim* s);
void rc_adjoint_step(rc_sim* s, long n);
void rc_adjoint_finish(rc_sim* s);
int rc_get_adjoint(const rc_sim* s, int field, int comp, double* out);
int rc_set_adjoint(rc_sim* s, int field, int comp, const double* in);
int rc_nthreads(int n); /* set OpenMP threads (0 = leave); returns the count in use */

#ifdef __cplusplus
}
#endif
#endif
