// Device side of the FDTD hot path for gfx950.
//
// The arithmetic of one cell lives in e_cell_3d / h_cell_3d (and the 2-D TM
// pair); two schedules call them:
//   * two passes per step (e_pass_3d, then h_pass_3d), in place;
//   * one fused pass per step (fused_3d): each workgroup owns a 254-column x
//     TY-row x kc-plane block, marches z, recomputes the new E of one halo
//     column, one halo row and one halo plane from the old state, and writes
//     E and H of its own cells to the "next" buffers (ping-pong), so every
//     cell's E and H are read once and written once per step.
// Both schedules feed the same inputs to the same functions, so they give
// bitwise-identical fields.  Fields are structure-of-arrays, x fastest, so a
// 64-lane wavefront touches 64 consecutive cells of one component.  E and H
// are stored directly (E = D/(eps_r eps0) folded into the update); D, P and J
// exist only inside the dispersive box.
//
// Reference arithmetic restated here (pederson/fdtd):
//   curl      include/YeeUpdates.hpp:363-405   (E-type fluxes take backward
//             differences of H, B-type forward differences of E)
//   CPML      include/PML.hpp:315-323          (flux += s dt psi G with the old
//             psi, then psi = B psi + C eps_IJK/dx diff)
//   material  include/DispersiveMaterials.hpp:354-359, 710-735, 979-998
//   TFSF      include/Source.hpp:506-553, interpolation 442-458
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

// Pointers that kernels read out of argument blocks in device memory are
// declared in the global address space for the device compile, so every
// access through them is a global load/store (not a flat one, which would
// also tie up the LDS counter and force full waits); the host compile sees
// plain pointers of the same size and layout.
#if defined(__HIP_DEVICE_COMPILE__)
#define FDTD_G __attribute__((address_space(1)))
#else
#define FDTD_G
#endif

namespace fdtdk {

// The cell arithmetic below turns off automatic FMA contraction and writes
// its fused multiply-adds explicitly: with contraction left to the compiler,
// an expression such as a*b - c*d may be fused on either product depending on
// the surrounding code, so schedules that inline the same cell function into
// different kernels could round differently.  Explicit FMAs keep every
// schedule bitwise equal.
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

constexpr int kMaxPoles = 3;
constexpr int kMaxSrc = 16;
constexpr int kMaxDft = 8;
constexpr int kMaxProbe = 32;

template <typename T>
struct MatDev {
  T inv_eps;  // 1/(eps_r eps0)
  T inv_mu;   // 1/(mu_r mu0)
  int np;     // poles; 0 = constant material
  T cJ[kMaxPoles], cE[kMaxPoles], cP[kMaxPoles];  // J = cJ J + cE E + cP P
  T cJm[kMaxPoles];                                 // 1 - cJ, formed in FP64
  // conductive pole: P = (P + cD D)/(1 + cD) with cD = dt sigma/(eps_r eps0)
  // (ConductiveUpdate::implicit_update, DispersiveMaterials.hpp:549-560); 0 for
  // Drude / Lorentz poles, whose P advances by dt J
  T cD[kMaxPoles];
  T cG;      // 1/(1 + cD[0]): the FP32 form of a lone conductive pole
  T eps_e0;  // eps_r eps0
  // magnetized Drude (3-D): 0 none, 1 a fixed bias field, 2 the cell's own B
  // (MagnetizedDrudeUpdate3D::implicit_update, DispersiveMaterials.hpp:1283-1340)
  int mag;
  // fluid Drude: omega_p and gamma per cell (from the electron density and
  // collision rate), its coefficients in StepArgs::fl
  int fluid;
  T v[3];  // fixed bias: v = dt (q_e/m_e) B
  T w2dt;  // (dt omega_p)^2 / dt
  T op;    // 1 + dt gamma + (dt omega_p)^2
  T vq;    // dt q_e/m_e: v = vq B for the cell's own B
  // magnetic side: poles on B, H, M, K, the electric side's updates with mu0
  // (K = mcJ K + mcE H + mcP M; conductive: M = (M + mcD B)/(1 + mcD))
  int mnp;
  T mcJ[kMaxPoles], mcE[kMaxPoles], mcP[kMaxPoles], mcJm[kMaxPoles], mcD[kMaxPoles];
  T mcG;    // 1/(1 + mcD[0])
  T mu_m0;  // mu_r mu0
};

struct SrcDev {
  long cell;
  int field, comp;
  const FDTD_G double* table;
  int n;
};

template <typename T>
struct DftDev {
  int r[6];
  int mask, nfreq, ncomp;
  long ncells;
  // [ncomp][nfreq][2][ncells], FP64 for every field precision: a sum over
  // thousands of steps keeps FP64 rounding, so the bins of an FP32 grid carry
  // the error of its fields only (read and written once per kDftBlock steps)
  FDTD_G double* acc;
  const FDTD_G double* phE;  // [nphase][nfreq][2]
  const FDTD_G double* phH;
  int nphase;
  long base;      // offset of this monitor's (frequency, cell) items in the DFT launch
  long cbase;     // offset of this monitor's cells in the sampling launch
  long sbase;     // offset of this monitor's (component, padded cell) items in the fold launch
  FDTD_G T* smp;  // [kDftBlock][ncomp][ncells]: samples of steps not yet folded in
};

// steps of field samples gathered before they are folded into the DFT
// accumulators (which are then read and written once per block)
constexpr int kDftBlock = 16;

struct ProbeDev {
  long cell;
  int field, comp;
};

// One single TFSF plane (make_tfsf_plane, include/Source.hpp:624-634): the
// cells [lo, hi] (inclusive; lo[axis] == hi[axis] is the plane's layer), the
// outward normal nrm along `axis`, and its own incident line.
constexpr int kMaxPlane = 6;
struct PlaneDev {
  int axis;
  int lo[3], hi[3];
  double nrm;
  double kvec[3], pol[3], polH[3], origin[3];
  int auxL;
  const FDTD_G double* incE;  // the line's Ez after each step's advance, rows of the chunk
  const FDTD_G double* incH;  // its Hy before the advance
};

template <typename T>
struct StepArgs {
  FDTD_G T* E[3];   // state read by the step
  FDTD_G T* H[3];
  FDTD_G T* Eo[3];  // state written by the step (the same arrays for in-place passes)
  FDTD_G T* Ho[3];
  const FDTD_G uint8_t* mat;
  const FDTD_G MatDev<T>* mats;
  const FDTD_G T* mie;  // 1/(eps_r eps0) per material id, negated for dispersive materials
  const FDTD_G T* mim;  // 1/(mu_r mu0) per material id
  // CPML: psi[ft][I][J] slab arrays (read / written) and per-axis profiles
  FDTD_G T* psi[2][3][3];
  FDTD_G T* psio[2][3][3];
  const FDTD_G T* cf[2][3];  // dt F / dx
  const FDTD_G T* cg[2][3];  // dt G
  const FDTD_G T* cb[2][3];  // B
  const FDTD_G T* cc[2][3];  // C / dx
  int slab_a[3], slab_b[3], slab_S[3];
  int nx, ny, nz;
  long sy, sz;
  // dispersive box (dense): [x0,x1)x[y0,y1)x[z0,z1), read / written
  int dbox[6];
  FDTD_G T* Dd[3];
  FDTD_G T* Pd[kMaxPoles][3];
  FDTD_G T* Jd[kMaxPoles][3];
  FDTD_G T* Ddo[3];
  FDTD_G T* Pdo[kMaxPoles][3];
  FDTD_G T* Jdo[kMaxPoles][3];
  // cells of materials with magnetic poles, over their bounding box (updated
  // in place by the general H passes): B, M_p, K_p
  int mbox[6];
  FDTD_G T* Bd[3];
  FDTD_G T* Md[kMaxPoles][3];
  FDTD_G T* Kd[kMaxPoles][3];
  // this block's own device address (set by the host; read by cold branches)
  const FDTD_G StepArgs* self;
  // fluid Drude coefficients per dispersive-box cell: {cJ, cE} of the Drude
  // recursion (TM, or no magnetization) or {w2dt, op} of the magnetized one
  const FDTD_G T* fl[2];
  T dt;
  // soft sources
  int nsrc;
  SrcDev src[kMaxSrc];
  // TFSF: incident rows of the current chunk, row s = step - step0
  int tfsf;
  int box[6];
  // boundary conditions, per [axis][side]: the layer whose new values are
  // copied into the ghost layer (-1: no condition), the address offset from
  // a source cell to its ghost, and per field type (E, H) the components
  // that are negated (bit c); bc_any: some face has a condition
  int bc_src[3][2];
  long bc_off[3][2];
  int bc_neg[2][3][2];
  int bc_any;
  double kvec[3], pol[3], polH[3], origin[3], dx;
  int auxL;
  const FDTD_G double* incE;  // aux Ez after each step's aux advance
  const FDTD_G double* incH;  // aux Hy before it
  // running DFT
  int ndft;
  long dft_work;   // sum over monitors of nfreq x cells
  long dft_cells;  // the sampler's items: per monitor, its cells padded to whole 256-thread blocks
  long dft_slots;  // the fold's items: per monitor, components x its cells padded to whole blocks
  long dft_done0;  // steps folded into the accumulators when this chunk started
  DftDev<T> dft[kMaxDft];
  // probes
  int nprobe;
  ProbeDev probe[kMaxProbe];
  FDTD_G double* trace;  // [step][nprobe]
  long step0;     // global step of row 0 of the incident rows
  // design box [x0,x1)x[y0,y1)x[z0,z1): per-cell 1/(eps_r eps0) replaces the
  // material table's value there (null when the grid has no design box)
  const FDTD_G T* dsg_ie;
  int dsg[6];
  // E increment of the design cells, one row [3][ndes] per step of
  // [hist_step0, hist_step0 + hist_rows) (null = not recorded)
  FDTD_G T* hist;
  long hist_step0, hist_rows;
  unsigned field_bytes, mat_bytes;  // padded array sizes for buffer addressing (0: too large)
  // read by the general instances only (kept last, so the offsets of the
  // fields above stay small): soft sources on E or H among src[], single
  // TFSF planes
  int nsrc_eh;
  int nplane;
  PlaneDev plane[kMaxPlane];
};

// Extra elements allocated after every field array and the id volume: the
// streamed pass loads whole rows of its last tiles, past the last cell, and
// the pad keeps those loads inside the allocation.
__host__ __device__ constexpr long field_pad(int nx) { return 8L * nx + 128; }

// local index of (i,j,k) in the design box, -1 outside
template <typename T>
__device__ __forceinline__ long design_index(const StepArgs<T>& a, int i, int j, int k) {
  const int di = i - a.dsg[0], dj = j - a.dsg[2], dk = k - a.dsg[4];
  const int wx = a.dsg[1] - a.dsg[0], wy = a.dsg[3] - a.dsg[2], wz = a.dsg[5] - a.dsg[4];
  if ((unsigned)di >= (unsigned)wx || (unsigned)dj >= (unsigned)wy || (unsigned)dk >= (unsigned)wz) return -1;
  return (long)di + (long)wx * ((long)dj + (long)wy * dk);
}

// 1/(eps_r eps0) of a cell: the design value inside the design box, else the material's
template <typename T, typename Tab>
__device__ __forceinline__ T cell_inv_eps(const StepArgs<T>& a, const Tab& tab, int mid, int i, int j, int k) {
  T ie = tab[mid];
  if (a.dsg_ie) {
    const long l = design_index(a, i, j, k);
    if (l >= 0) ie = a.dsg_ie[l];
  }
  return ie;
}

__device__ __forceinline__ int slab_index(int idx, int a, int b) { return idx < a ? idx : (idx >= b ? a + idx - b : -1); }

// slab cell offset for axis J (x fastest within the slab array)
template <typename T>
__device__ __forceinline__ long slab_off(const StepArgs<T>& a, int J, int s, int i, int j, int k) {
  if (J == 0) return (long)s + (long)a.slab_S[0] * ((long)j + (long)a.ny * k);
  if (J == 1) return (long)i + (long)a.nx * ((long)s + (long)a.slab_S[1] * k);
  return (long)i + (long)a.nx * ((long)j + (long)a.ny * s);
}

// Aux1D::interpolateE / interpolateH (include/Source.hpp:442-458) on a row of
// the incident table.
template <typename Row>
__device__ __forceinline__ double interp_row(Row row, int L, double loc, double dx, bool h) {
#pragma clang fp contract(off)
  double nd = (double)L / 2.0 + loc / dx;
  if (h) nd -= 0.5;
  long n = (long)nd;
  if (n < 0) n = 0;
  if (n > L - 2) n = L - 2;
  double frac = nd - (double)n;
  return fmadd(frac, row[n + 1], (1.0 - frac) * row[n]);
}

// Levi-Civita of distinct axes
__device__ __forceinline__ double lc3(int I, int J) {
  int K = 3 - I - J;
  return ((I == 0 && J == 1 && K == 2) || (I == 1 && J == 2 && K == 0) || (I == 2 && J == 0 && K == 1)) ? 1.0 : -1.0;
}

// TFSF correction of one face cell (TFSFPlane::field_update, Source.hpp:542-553).
// ft 0: D fluxes from aux H with polH; ft 1: B fluxes from aux E with pol.
// The total-field region along an axis is the open interval (lo-1/2, hi);
// a component sits at offset 0 or -1/2 on each axis.
// Cold paths (TFSF faces, sources, dispersive cells, DFT monitors).  Inlined:
// out-of-line calls measured more registers (call clobbers) plus scratch for
// the argument arrays.  Kernel arguments are read through a pointer to device
// memory, so inlined code never copies the large argument struct.
#ifndef FDTD_COLD
#define FDTD_COLD __forceinline__
#endif

// The argument block as a cold branch should read it: through `self`, the
// block's own device address stored in it by the host.  A pointer read from
// memory is not known to be dereferenceable, so the branch's loads through it
// cannot be hoisted out of the branch into scalar registers for the whole
// kernel (where, with the hot path's values, they would spill into VGPR
// lanes).  Blocks without `self` are read directly.  The view is typed in the
// global address space so its loads are global, not flat, loads.
template <typename T>
__device__ __forceinline__ const FDTD_G StepArgs<T>& cold_args(const StepArgs<T>& a) {
  const FDTD_G StepArgs<T>* p = a.self ? a.self : (const FDTD_G StepArgs<T>*)&a;
  return *p;
}

// COLD: read the TFSF parameters through the cold view of the argument block
// (the E pass of grids with dispersive materials, whose hoisted scalars would
// otherwise spill: 77 -> 69 VGPRs, 6 -> 7 waves); the plain E pass and the H
// pass keep them in scalar registers (the cold view measured slower there).  (The cold view is read as
// a generic reference here: typed in the global address space it leaves the
// E pass with more registers, measured.)
// row: the step's incident row (global memory, or its copy in LDS); fn: the
// cell's curl coefficients dt F/dx per axis for this field type (null: read
// from the argument block, which keeps them out of registers until needed).
// J0: the first face axis taken (the time-blocked pass adds the x faces' terms itself)
template <typename T, int NDIM, bool COLD = false, typename Row = const FDTD_G double*, int J0 = 0>
__device__ FDTD_COLD void tfsf_cell(const StepArgs<T>& a0, Row row, int ft, int i, int j, int k, const bool actF[3],
                                    const bool actC[3], const T fn[3], T delta[3]) {
#pragma clang fp contract(off)
  const StepArgs<T>& a = COLD ? (const StepArgs<T>&)cold_args(a0) : a0;
  const int ci[3] = {i, j, k};
  // loops fully unrolled: every index of delta[] and ci[] is a constant, so
  // neither array needs memory (a dynamically indexed one would be placed in
  // LDS or scratch)
#pragma unroll
  for (int J = J0; J < NDIM; J++) {
    int side = (ci[J] == a.box[2 * J]) ? 0 : ((ci[J] == a.box[2 * J + 1]) ? 1 : -1);
    if (side < 0) continue;
    const double nrm = side == 0 ? -1.0 : 1.0;
#pragma unroll
    for (int I = 0; I < 3; I++) {
      if (I == J || !actF[I]) continue;
      const int K = 3 - I - J;
      if (!actC[K]) continue;
      bool inside = true;
#pragma unroll
      for (int Tn = 0; Tn < NDIM; Tn++) {
        if (Tn == J) continue;
        bool half = ft == 0 ? (Tn != I) : (Tn == I);
        int lo = half ? a.box[2 * Tn] + 1 : a.box[2 * Tn];
        int hi = half ? a.box[2 * Tn + 1] : a.box[2 * Tn + 1] - 1;
        if (ci[Tn] < lo || ci[Tn] > hi) inside = false;
      }
      if (!inside) continue;
      double proj = 0.0;
#pragma unroll
      for (int dd = 0; dd < NDIM; dd++) {
        double off = ft == 0 ? (dd == K ? -0.5 : 0.0) : (dd != K ? -0.5 : 0.0);
        double loc = fmadd(off, a.dx, ((double)ci[dd] - a.origin[dd]) * a.dx);
        proj = fmadd(a.kvec[dd], loc, proj);
      }
      double inc = interp_row(row, a.auxL, proj, a.dx, ft == 0);
      double pv = ft == 0 ? a.polH[K] : a.pol[K];
      double f = fn ? (double)fn[J] : (double)a.cf[ft][J][ci[J]];
      double term = lc3(I, J) * f * nrm * inc * pv;
      delta[I] += (T)(ft == 0 ? term : -term);
    }
  }
}

// Single TFSF planes through cell (i,j,k) (TFSFPlane::updateD / updateB,
// include/Source.hpp:595-620): every plane holding the cell adds, for each
// flux I other than its normal axis J, eps_IJK (dt F_J/dx) n_J inc_K pol_K
// to the cell's delta (with the curl's sign: + for D, - for B), inc_K the
// plane's own line at k.(r + off_K dx), r = (cell - origin) dx.  ft 0: D from
// the line's H with polH; 1: B from its E with pol.  actF / actC: the mode's
// flux and curl components (bit per axis).
template <typename T, int NDIM>
__device__ FDTD_COLD void plane_cells(const StepArgs<T>& a0, long step, int ft, int i, int j, int k, int actF,
                                      int actC, T delta[3]) {
#pragma clang fp contract(off)
  const StepArgs<T>& a = cold_args(a0);
  const long s = step - a.step0;
  for (int p = 0; p < a.nplane; p++) {
    const PlaneDev& P = a.plane[p];
    if (i < P.lo[0] || i > P.hi[0] || j < P.lo[1] || j > P.hi[1] || k < P.lo[2] || k > P.hi[2]) continue;
    const int J = P.axis;
    const FDTD_G double* row = (ft == 0 ? P.incH : P.incE) + s * P.auxL;
    const double f = (double)a.cf[ft][J][J == 0 ? i : (J == 1 ? j : k)];
#pragma unroll
    for (int I = 0; I < 3; I++) {
      if (I == J || !((actF >> I) & 1)) continue;
      const int K = 3 - I - J;
      if (!((actC >> K) & 1)) continue;
      double proj = 0.0;
#pragma unroll
      for (int dd = 0; dd < NDIM; dd++) {
        const int cd = dd == 0 ? i : (dd == 1 ? j : k);
        const double off = ft == 0 ? (dd == K ? -0.5 : 0.0) : (dd != K ? -0.5 : 0.0);
        const double loc = fmadd(off, a.dx, ((double)cd - P.origin[dd]) * a.dx);
        proj = fmadd(P.kvec[dd], loc, proj);
      }
      const double inc = interp_row(row, P.auxL, proj, a.dx, ft == 0);
      const double pv = ft == 0 ? P.polH[K] : P.pol[K];
      const double term = lc3(I, J) * f * P.nrm * inc * pv;
      const T v = (T)(ft == 0 ? term : -term);
      if (I == 0)
        delta[0] += v;
      else if (I == 1)
        delta[1] += v;
      else
        delta[2] += v;
    }
  }
}

template <typename T>
__device__ __forceinline__ T pick3(T v0, T v1, T v2, int c) {
  return c == 0 ? v0 : (c == 1 ? v1 : v2);
}

template <typename T, bool COLD = true>
__device__ FDTD_COLD void add_sources(const StepArgs<T>& a0, long step, int field, long c, T d[3]) {
  const StepArgs<T>& a = COLD ? (const StepArgs<T>&)cold_args(a0) : a0;
  for (int s = 0; s < a.nsrc; s++) {
    const SrcDev& q = a.src[s];
    if (q.cell == c && q.field == field && step < q.n) {
      const T v = (T)q.table[step];
      if (q.comp == 0)
        d[0] += v;
      else if (q.comp == 1)
        d[1] += v;
      else
        d[2] += v;
    }
  }
}

// Soft sources on the derived fields, E (field 1) or H (field 3): SoftSource
// on E / H, which the reference's loops apply after the material step
// (include/Source.hpp:209-229; Aux1D places its own so, 421-425), so that the
// next material step, recomputing the field from D or B, drops the value
// again.  E and H are stored here without D and B (outside dispersive
// cells), so the stored value carries the last step's source value: with
// `back` the update first takes table[step - 1] off the cell's old value;
// without, it adds table[step] to the new one.
template <typename T>
__device__ FDTD_COLD void soft_eh(const StepArgs<T>& a0, long step, int field, long c, bool back, T v[3]) {
  const StepArgs<T>& a = cold_args(a0);
  const long n = back ? step - 1 : step;
  for (int s = 0; s < a.nsrc; s++) {
    const SrcDev& q = a.src[s];
    if (q.cell == c && q.field == field && n >= 0 && n < q.n) {
      const T x = back ? -(T)q.table[n] : (T)q.table[n];
      if (q.comp == 0)
        v[0] += x;
      else if (q.comp == 1)
        v[1] += x;
      else
        v[2] += x;
    }
  }
}

// Dispersive cell, one component: D += delta; P_p += dt J_p;
// E = (D - sum_p P_p)/(eps_r eps0); J_p = cJ J_p + cE E + cP P_p
// (DrudeUpdate / LorentzUpdate implicit_update, DispersiveMaterials.hpp:710-735, 979-998).
template <typename T, int I>
__device__ __forceinline__ T ade_component(const StepArgs<T>& a, const MatDev<T>& m, long bq, T d, T Eold, bool own) {
#pragma clang fp contract(off)
  const T D = a.Dd[I][bq] + d;
  if (sizeof(T) == 4 && m.np == 1 && m.cD[0] != T(0)) {
    // FP32, a lone conductive pole: E_{n+1} = (E_n + d/(eps_r eps0))/(1 + f),
    // the same update without the cancellation of D - P (P -> D in a good
    // conductor, which in FP32 leaves E at a few digits); P = D - eps_r eps0 E
    const T Ev = fmadd(d, m.inv_eps, Eold) * m.cG;
    if (own) {
      a.Ddo[I][bq] = D;
      a.Pdo[0][I][bq] = fmadd(-m.eps_e0, Ev, D);
      a.Jdo[0][I][bq] = a.Jd[0][I][bq];
    }
    return Ev;
  }
  T num = D;
  T Pv[kMaxPoles], Jv[kMaxPoles];
#pragma unroll
  for (int p = 0; p < kMaxPoles; p++) {
    if (p < m.np) {
      Jv[p] = a.Jd[p][I][bq];
      if (m.cD[p] != T(0))
        Pv[p] = fmadd(m.cD[p], D, a.Pd[p][I][bq]) / (T(1) + m.cD[p]);
      else
        Pv[p] = fmadd(a.dt, Jv[p], a.Pd[p][I][bq]);
      num -= Pv[p];
    }
  }
  T Ev = num * m.inv_eps;
  if (sizeof(T) == 4 && m.cD[0] == T(0) && m.cD[1] == T(0) && m.cD[2] == T(0)) {
    // FP32, Drude / Lorentz poles: E_{n+1} = E_n + (d - dt sum_p J_p)/(eps_r eps0),
    // equal to (D - sum_p P_p)/(eps_r eps0) (each P_p advances by dt J_p)
    // without forming D - sum P, which inside a metal cancels to a small E and
    // leaves it a few digits in FP32 (FP64 keeps the reference's form)
    T inc = d;
#pragma unroll
    for (int p = 0; p < kMaxPoles; p++)
      if (p < m.np) inc = fmadd(-a.dt, Jv[p], inc);
    Ev = fmadd(inc, m.inv_eps, Eold);
  }
  if (own) {
    a.Ddo[I][bq] = D;
#pragma unroll
    for (int p = 0; p < kMaxPoles; p++) {
      if (p < m.np) {
        a.Pdo[p][I][bq] = Pv[p];
        // FP32: cJ J as J - (1 - cJ) J.  cJ is within a few 1e-3 of 1 and its FP32
        // rounding (~6e-8) compounds over the steps (cJ^n); 1 - cJ is stored to
        // FP32's relative precision, so the pole's decay stays that of FP64
        const T jj = sizeof(T) == 4 ? fmadd(-m.cJm[p], Jv[p], Jv[p]) : m.cJ[p] * Jv[p];
        a.Jdo[p][I][bq] = fmadd(m.cP[p], Pv[p], fmadd(m.cE[p], Ev, jj));
      }
    }
  }
  return Ev;
}

template <typename T>
__device__ __forceinline__ long dbox_index(const StepArgs<T>& a, int i, int j, int k) {
  return (long)(i - a.dbox[0]) +
         (long)(a.dbox[1] - a.dbox[0]) * ((long)(j - a.dbox[2]) + (long)(a.dbox[3] - a.dbox[2]) * (long)(k - a.dbox[4]));
}

template <typename T>
__device__ __forceinline__ long mbox_index(const StepArgs<T>& a, int i, int j, int k) {
  return (long)(i - a.mbox[0]) +
         (long)(a.mbox[1] - a.mbox[0]) * ((long)(j - a.mbox[2]) + (long)(a.mbox[3] - a.mbox[2]) * (long)(k - a.mbox[4]));
}

// Magnetic-side pole cell, one component (the reference's pole functors on
// FieldType::Magnetic, which run the electric updates on B, H, M, K with mu0):
// B += d; M_p += dt K_p (conductive: (M_p + mcD B)/(1 + mcD));
// H = (B - sum_p M_p)/(mu_r mu0); K_p = mcJ K_p + mcE H + mcP M_p.  FP32 takes
// the same forms as ade_component (H from its old value where B - sum M
// cancels).  In place: the general H passes own their cells.
template <typename T, int I>
__device__ FDTD_COLD T mag_component(const StepArgs<T>& a0, long c, long bq, T d, T Hold) {
#pragma clang fp contract(off)
  const StepArgs<T>& a = (const StepArgs<T>&)cold_args(a0);
  const MatDev<T>& m = a.mats[a.mat[c]];
  const T B = a.Bd[I][bq] + d;
  if (sizeof(T) == 4 && m.mnp == 1 && m.mcD[0] != T(0)) {
    const T Hv = fmadd(d, m.inv_mu, Hold) * m.mcG;
    a.Bd[I][bq] = B;
    a.Md[0][I][bq] = fmadd(-m.mu_m0, Hv, B);
    return Hv;
  }
  T num = B;
  T Mv[kMaxPoles], Kv[kMaxPoles];
#pragma unroll
  for (int p = 0; p < kMaxPoles; p++) {
    if (p < m.mnp) {
      Kv[p] = a.Kd[p][I][bq];
      if (m.mcD[p] != T(0))
        Mv[p] = fmadd(m.mcD[p], B, a.Md[p][I][bq]) / (T(1) + m.mcD[p]);
      else
        Mv[p] = fmadd(a.dt, Kv[p], a.Md[p][I][bq]);
      num -= Mv[p];
    }
  }
  T Hv = num * m.inv_mu;
  if (sizeof(T) == 4 && m.mcD[0] == T(0) && m.mcD[1] == T(0) && m.mcD[2] == T(0)) {
    T inc = d;
#pragma unroll
    for (int p = 0; p < kMaxPoles; p++)
      if (p < m.mnp) inc = fmadd(-a.dt, Kv[p], inc);
    Hv = fmadd(inc, m.inv_mu, Hold);
  }
  a.Bd[I][bq] = B;
#pragma unroll
  for (int p = 0; p < kMaxPoles; p++) {
    if (p < m.mnp) {
      a.Md[p][I][bq] = Mv[p];
      const T kk = sizeof(T) == 4 ? fmadd(-m.mcJm[p], Kv[p], Kv[p]) : m.mcJ[p] * Kv[p];
      a.Kd[p][I][bq] = fmadd(m.mcP[p], Mv[p], fmadd(m.mcE[p], Hv, kk));
    }
  }
  return Hv;
}

// Fluid Drude, one component, with the cell's own coefficients cJ, cE
// (FluidDrudeStorage: omega_p = K sqrt(ne), gamma = nu_m, through
// DrudeUpdate::implicit_update, DispersiveMaterials.hpp:979-998, 1160-1181):
// P += dt J; E = (D - P)/eps0; J = cJ J + cE E.
template <typename T, int I>
__device__ __forceinline__ T fluid_component(const StepArgs<T>& a, const MatDev<T>& m, long bq, T d, bool own) {
  const T D = a.Dd[I][bq] + d;
  const T J = a.Jd[0][I][bq];
  const T P = a.Pd[0][I][bq] + a.dt * J;
  const T Ev = (D - P) * m.inv_eps;
  if (own) {
    a.Ddo[I][bq] = D;
    a.Pdo[0][I][bq] = P;
    a.Jdo[0][I][bq] = a.fl[0][bq] * J + a.fl[1][bq] * Ev;
  }
  return Ev;
}

// Magnetized Drude cell (3-D): the implicit Euler system couples the three
// components through v = dt (q_e/m_e) B.  With r = J + (wp^2/dt)(D - P) and
// M = A/denom (A = op^2 I + v v^T + op [v]x, denom = op (op^2 + |v|^2)):
// J' = M r, P' = P + dt J' (the reference's P' = P + M (dt r), the same
// vector), E = (D - P')/(eps_r eps0)
// (MagnetizedDrudeUpdate3D::implicit_update, DispersiveMaterials.hpp:1283-1340).
// B is the fixed bias (MagnetizedDrudeStorage) or the cell's own field
// B = mu_r mu0 H (the self-magnetized storages, 1816-1943); wp and gamma are
// the material's or, for the fluid forms, the cell's own (1728-1893).
// A fluid material without magnetization runs fluid_component.
template <typename T>
__device__ FDTD_COLD void extended_cell(const StepArgs<T>& a, const MatDev<T>& m, long bq, const T d[3],
                                        const T hown[3], bool own, T out[3]) {
  if (!m.mag) {
    out[0] = fluid_component<T, 0>(a, m, bq, d[0], own);
    out[1] = fluid_component<T, 1>(a, m, bq, d[1], own);
    out[2] = fluid_component<T, 2>(a, m, bq, d[2], own);
    return;
  }
  const T w2dt = m.fluid ? a.fl[0][bq] : m.w2dt;
  const T op = m.fluid ? a.fl[1][bq] : m.op;
  T P[3], s[3], r[3];  // s = D - P
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const T D = a.Dd[c][bq] + d[c];
    if (own) a.Ddo[c][bq] = D;
    P[c] = a.Pd[0][c][bq];
    s[c] = D - P[c];
    r[c] = a.Jd[0][c][bq] + w2dt * s[c];
  }
  T v[3];
#pragma unroll
  for (int c = 0; c < 3; c++) v[c] = m.mag == 1 ? m.v[c] : m.vq * (hown[c] / m.inv_mu);
  const T inv = T(1) / (op * (op * op + (v[0] * v[0] + v[1] * v[1] + v[2] * v[2])));
#pragma unroll
  for (int c = 0; c < 3; c++) {
    // row c of A = op^2 I + v v^T + op [v]x, [v]x = ((0,-vz,vy),(vz,0,-vx),(-vy,vx,0))
    const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
    T row[3];
    row[c] = op * op + v[c] * v[c];
    row[c1] = -v[c2] * op + v[c] * v[c1];
    row[c2] = v[c1] * op + v[c] * v[c2];
    const T Jn = (row[0] * r[0] + row[1] * r[1] + row[2] * r[2]) * inv;
    out[c] = (s[c] - a.dt * Jn) * m.inv_eps;
    if (own) {
      a.Pdo[0][c][bq] = P[c] + a.dt * Jn;
      a.Jdo[0][c][bq] = Jn;
    }
  }
}

// dispersive cell: the three components (hown: the cell's own H, read by a
// self-magnetized Drude material; MAG: the instance that carries the
// magnetized and fluid Drude code)
template <typename T, bool MAG>
__device__ FDTD_COLD void ade_cell(const StepArgs<T>& a, int mid, int i, int j, int k, const T d[3], const T eold[3],
                                   const T hown[3], bool own, T out[3]) {
  const MatDev<T>& m = a.mats[mid];
  const long bq = dbox_index(a, i, j, k);
  if (MAG && (m.mag || m.fluid)) {
    extended_cell(a, m, bq, d, hown, own, out);
    return;
  }
  out[0] = ade_component<T, 0>(a, m, bq, d[0], eold[0], own);
  out[1] = ade_component<T, 1>(a, m, bq, d[1], eold[1], own);
  out[2] = ade_component<T, 2>(a, m, bq, d[2], eold[2], own);
}

// Inputs of a cell update that a caller may gather ahead of it (the streamed
// pass does, a plane early): per slab axis J the cell's two CPML psi of its
// field type (components I != J in increasing order) and the axis
// coefficients dt G, B and C/dx at the cell's coordinate.
// One soft source's value at the current step, staged where the update
// reads it (field 0: D, 2: B; the value already converted to the field type).
template <typename T>
struct SrcVal {
  long cell;
  int field, comp;
  T v;
};

template <typename T>
struct CellPre {
  mutable T psi[3][2];  // in (PRE bit 0) and, with PRE bit 2, out
  T g[3], b[3], c[3];
  const SrcVal<T>* src;  // the step's source values (null: none)
  int nsrc;
};

// the staged form of add_sources: the same additions in the same order
template <typename T>
__device__ __forceinline__ void add_staged_sources(const CellPre<T>& pre, int field, long c, T d[3]) {
  for (int s = 0; s < pre.nsrc; s++) {
    const SrcVal<T>& q = pre.src[s];
    if (q.cell == c && q.field == field) {
      if (q.comp == 0)
        d[0] += q.v;
      else if (q.comp == 1)
        d[1] += q.v;
      else
        d[2] += q.v;
    }
  }
}

// H values around a cell that the E update reads (backward neighbours)
template <typename T>
struct HStencil {
  T hx, hy, hz, hz_ym, hz_xm, hy_xm, hy_zm, hx_ym, hx_zm;
};
// E values around a cell that the H update reads (forward neighbours)
template <typename T>
struct EStencil {
  T ex, ey, ez, ez_yp, ez_xp, ey_xp, ey_zp, ex_yp, ex_zp;
};

// ---------------------------------------------------------------- E of one cell
// New E of interior cell (i,j,k).  psi and the dispersive state are read from
// the current buffers and, when `own`, written to the next buffers.
// DISP = 0: the grid has no dispersive material (the ADE code is left out of
// that instance, which keeps its register count down); 1: pole materials;
// 2: also magnetized and fluid Drude materials.
// IP: the in-place passes, whose psi is read and written in the same arrays
// (one set of pointers instead of two: fewer scalar registers).
// PRE: bit 0: psi and the slab coefficients come from *pre (gathered by the
// caller ahead of the cell's arithmetic) instead of memory; bit 1: so do the
// step's source values (staged), and the TFSF correction takes the curl
// coefficients from the caller; bit 2: the new psi goes back into pre->psi
// instead of memory (the time-blocked pass keeps it in registers).  Row: the incident row's pointer type (global
// or LDS).
// GEN: the general instance, which also carries E / H soft sources and single
// TFSF planes (code the other instances leave out, keeping their registers).
template <typename T, int DISP = 1, bool IP = false, int PRE = 0, typename Row = const FDTD_G double*,
          bool GEN = false>
__device__ __forceinline__ void e_cell_3d(const StepArgs<T>& a, long step, Row incH, int i, int j, int k, long c,
                                          const HStencil<T>& h, T ex, T ey, T ez, int mid, T ie, T fx, T fy, T fz,
                                          bool own, T out[3], const CellPre<T>* pre = nullptr) {
#pragma clang fp contract(off)
  if (GEN && a.nsrc_eh) {  // the old E without the last step's E source value
    T v[3] = {ex, ey, ez};
    soft_eh(a, step, 1, c, true, v);
    ex = v[0];
    ey = v[1];
    ez = v[2];
  }
  const T dxy = h.hz - h.hz_ym;  // Ex, J=y, +1
  const T dxz = h.hy - h.hy_zm;  // Ex, J=z, -1
  const T dyz = h.hx - h.hx_zm;  // Ey, J=z, +1
  const T dyx = h.hz - h.hz_xm;  // Ey, J=x, -1
  const T dzx = h.hy - h.hy_xm;  // Ez, J=x, +1
  const T dzy = h.hx - h.hx_ym;  // Ez, J=y, -1
  T d[3];
  d[0] = fmadd(fy, dxy, -(fz * dxz));
  d[1] = fmadd(fz, dyz, -(fx * dyx));
  d[2] = fmadd(fx, dzx, -(fy * dzy));
  if (PRE & 2)
    add_staged_sources(*pre, 0, c, d);
  else if (a.nsrc)
    add_sources<T, true>(a, step, 0, c, d);
  const int sxi = slab_index(i, a.slab_a[0], a.slab_b[0]);
  if (sxi >= 0) {
    const long o = slab_off(a, 0, sxi, i, j, k);
    const T g = (PRE & 1) ? pre->g[0] : a.cg[0][0][i], b = (PRE & 1) ? pre->b[0] : a.cb[0][0][i],
            cc = (PRE & 1) ? pre->c[0] : a.cc[0][0][i];
    const T p1 = (PRE & 1) ? pre->psi[0][0] : a.psi[0][1][0][o], p2 = (PRE & 1) ? pre->psi[0][1] : a.psi[0][2][0][o];
    d[1] = fmadd(g, p1, d[1]);
    d[2] = fmadd(g, p2, d[2]);
    if (PRE & 4) {
      pre->psi[0][0] = fmadd(b, p1, -(cc * dyx));
      pre->psi[0][1] = fmadd(b, p2, cc * dzx);
    } else if (own) {
      (IP ? a.psi : a.psio)[0][1][0][o] = fmadd(b, p1, -(cc * dyx));
      (IP ? a.psi : a.psio)[0][2][0][o] = fmadd(b, p2, cc * dzx);
    }
  }
  const int syi = slab_index(j, a.slab_a[1], a.slab_b[1]);
  if (syi >= 0) {
    const long o = slab_off(a, 1, syi, i, j, k);
    const T g = (PRE & 1) ? pre->g[1] : a.cg[0][1][j], b = (PRE & 1) ? pre->b[1] : a.cb[0][1][j],
            cc = (PRE & 1) ? pre->c[1] : a.cc[0][1][j];
    const T p0 = (PRE & 1) ? pre->psi[1][0] : a.psi[0][0][1][o], p2 = (PRE & 1) ? pre->psi[1][1] : a.psi[0][2][1][o];
    d[0] = fmadd(g, p0, d[0]);
    d[2] = fmadd(g, p2, d[2]);
    if (PRE & 4) {
      pre->psi[1][0] = fmadd(b, p0, cc * dxy);
      pre->psi[1][1] = fmadd(b, p2, -(cc * dzy));
    } else if (own) {
      (IP ? a.psi : a.psio)[0][0][1][o] = fmadd(b, p0, cc * dxy);
      (IP ? a.psi : a.psio)[0][2][1][o] = fmadd(b, p2, -(cc * dzy));
    }
  }
  const int szi = slab_index(k, a.slab_a[2], a.slab_b[2]);
  if (szi >= 0) {
    const long o = slab_off(a, 2, szi, i, j, k);
    const T g = (PRE & 1) ? pre->g[2] : a.cg[0][2][k], b = (PRE & 1) ? pre->b[2] : a.cb[0][2][k],
            cc = (PRE & 1) ? pre->c[2] : a.cc[0][2][k];
    const T p0 = (PRE & 1) ? pre->psi[2][0] : a.psi[0][0][2][o], p1 = (PRE & 1) ? pre->psi[2][1] : a.psi[0][1][2][o];
    d[0] = fmadd(g, p0, d[0]);
    d[1] = fmadd(g, p1, d[1]);
    if (PRE & 4) {
      pre->psi[2][0] = fmadd(b, p0, -(cc * dxz));
      pre->psi[2][1] = fmadd(b, p1, cc * dyz);
    } else if (own) {
      (IP ? a.psi : a.psio)[0][0][2][o] = fmadd(b, p0, -(cc * dxz));
      (IP ? a.psi : a.psio)[0][1][2][o] = fmadd(b, p1, cc * dyz);
    }
  }
  if (a.tfsf && (i == a.box[0] || i == a.box[1] || j == a.box[2] || j == a.box[3] || k == a.box[4] || k == a.box[5])) {
    const bool act[3] = {true, true, true};
    const T fn[3] = {fx, fy, fz};
    tfsf_cell<T, 3, (DISP > 0)>(a, incH, 0, i, j, k, act, act, (PRE & 2) ? fn : nullptr, d);
  }
  if (GEN && a.nplane) plane_cells<T, 3>(a, step, 0, i, j, k, 7, 7, d);
  if (ie > T(0)) {
    out[0] = fmadd(d[0], ie, ex);
    out[1] = fmadd(d[1], ie, ey);
    out[2] = fmadd(d[2], ie, ez);
  } else if (DISP && ie < T(0) && a.Dd[0]) {
    const T eold[3] = {ex, ey, ez};
    T hown[3] = {h.hx, h.hy, h.hz};
    if (DISP == 2 && GEN && a.nsrc_eh) soft_eh(a, step, 3, c, true, hown);  // B = mu H without the H source
    ade_cell<T, DISP == 2>(a, mid, i, j, k, d, eold, hown, own, out);
  } else {
    out[0] = ex;
    out[1] = ey;
    out[2] = ez;
  }
  if (GEN && a.nsrc_eh) soft_eh(a, step, 1, c, false, out);
}

// ---------------------------------------------------------------- H of one cell
template <typename T, bool IP = false, int PRE = 0, typename Row = const FDTD_G double*, bool GEN = false>
__device__ __forceinline__ void h_cell_3d(const StepArgs<T>& a, long step, Row incE, int i, int j, int k, long c,
                                          const EStencil<T>& e, T hx, T hy, T hz, T im, T fx, T fy, T fz, T out[3],
                                          const CellPre<T>* pre = nullptr) {
#pragma clang fp contract(off)
  if (GEN && a.nsrc_eh) {  // the old H without the last step's H source value
    T v[3] = {hx, hy, hz};
    soft_eh(a, step, 3, c, true, v);
    hx = v[0];
    hy = v[1];
    hz = v[2];
  }
  const T dxy = e.ez_yp - e.ez;  // Hx, J=y, +1
  const T dxz = e.ey_zp - e.ey;  // Hx, J=z, -1
  const T dyz = e.ex_zp - e.ex;  // Hy, J=z, +1
  const T dyx = e.ez_xp - e.ez;  // Hy, J=x, -1
  const T dzx = e.ey_xp - e.ey;  // Hz, J=x, +1
  const T dzy = e.ex_yp - e.ex;  // Hz, J=y, -1
  T d[3];
  d[0] = fmadd(fz, dxz, -(fy * dxy));
  d[1] = fmadd(fx, dyx, -(fz * dyz));
  d[2] = fmadd(fy, dzy, -(fx * dzx));
  if (PRE & 2)
    add_staged_sources(*pre, 2, c, d);
  else if (a.nsrc)
    add_sources(a, step, 2, c, d);
  const int sxi = slab_index(i, a.slab_a[0], a.slab_b[0]);
  if (sxi >= 0) {
    const long o = slab_off(a, 0, sxi, i, j, k);
    const T g = (PRE & 1) ? pre->g[0] : a.cg[1][0][i], b = (PRE & 1) ? pre->b[0] : a.cb[1][0][i],
            cc = (PRE & 1) ? pre->c[0] : a.cc[1][0][i];
    const T p1 = (PRE & 1) ? pre->psi[0][0] : a.psi[1][1][0][o], p2 = (PRE & 1) ? pre->psi[0][1] : a.psi[1][2][0][o];
    d[1] = fmadd(-g, p1, d[1]);
    d[2] = fmadd(-g, p2, d[2]);
    if (PRE & 4) {
      pre->psi[0][0] = fmadd(b, p1, -(cc * dyx));
      pre->psi[0][1] = fmadd(b, p2, cc * dzx);
    } else {
      (IP ? a.psi : a.psio)[1][1][0][o] = fmadd(b, p1, -(cc * dyx));
      (IP ? a.psi : a.psio)[1][2][0][o] = fmadd(b, p2, cc * dzx);
    }
  }
  const int syi = slab_index(j, a.slab_a[1], a.slab_b[1]);
  if (syi >= 0) {
    const long o = slab_off(a, 1, syi, i, j, k);
    const T g = (PRE & 1) ? pre->g[1] : a.cg[1][1][j], b = (PRE & 1) ? pre->b[1] : a.cb[1][1][j],
            cc = (PRE & 1) ? pre->c[1] : a.cc[1][1][j];
    const T p0 = (PRE & 1) ? pre->psi[1][0] : a.psi[1][0][1][o], p2 = (PRE & 1) ? pre->psi[1][1] : a.psi[1][2][1][o];
    d[0] = fmadd(-g, p0, d[0]);
    d[2] = fmadd(-g, p2, d[2]);
    if (PRE & 4) {
      pre->psi[1][0] = fmadd(b, p0, cc * dxy);
      pre->psi[1][1] = fmadd(b, p2, -(cc * dzy));
    } else {
      (IP ? a.psi : a.psio)[1][0][1][o] = fmadd(b, p0, cc * dxy);
      (IP ? a.psi : a.psio)[1][2][1][o] = fmadd(b, p2, -(cc * dzy));
    }
  }
  const int szi = slab_index(k, a.slab_a[2], a.slab_b[2]);
  if (szi >= 0) {
    const long o = slab_off(a, 2, szi, i, j, k);
    const T g = (PRE & 1) ? pre->g[2] : a.cg[1][2][k], b = (PRE & 1) ? pre->b[2] : a.cb[1][2][k],
            cc = (PRE & 1) ? pre->c[2] : a.cc[1][2][k];
    const T p0 = (PRE & 1) ? pre->psi[2][0] : a.psi[1][0][2][o], p1 = (PRE & 1) ? pre->psi[2][1] : a.psi[1][1][2][o];
    d[0] = fmadd(-g, p0, d[0]);
    d[1] = fmadd(-g, p1, d[1]);
    if (PRE & 4) {
      pre->psi[2][0] = fmadd(b, p0, -(cc * dxz));
      pre->psi[2][1] = fmadd(b, p1, cc * dyz);
    } else {
      (IP ? a.psi : a.psio)[1][0][2][o] = fmadd(b, p0, -(cc * dxz));
      (IP ? a.psi : a.psio)[1][1][2][o] = fmadd(b, p1, cc * dyz);
    }
  }
  if (a.tfsf && (i == a.box[0] || i == a.box[1] || j == a.box[2] || j == a.box[3] || k == a.box[4] || k == a.box[5])) {
    const bool act[3] = {true, true, true};
    const T fn[3] = {fx, fy, fz};
    tfsf_cell<T, 3>(a, incE, 1, i, j, k, act, act, (PRE & 2) ? fn : nullptr, d);
  }
  if (GEN && a.nplane) plane_cells<T, 3>(a, step, 1, i, j, k, 7, 7, d);
  if (GEN && im < T(0)) {  // magnetic poles (the general instance only)
    const long bq = mbox_index(a, i, j, k);
    out[0] = mag_component<T, 0>(a, c, bq, d[0], hx);
    out[1] = mag_component<T, 1>(a, c, bq, d[1], hy);
    out[2] = mag_component<T, 2>(a, c, bq, d[2], hz);
  } else {
    out[0] = fmadd(d[0], im, hx);
    out[1] = fmadd(d[1], im, hy);
    out[2] = fmadd(d[2], im, hz);
  }
  if (GEN && a.nsrc_eh) soft_eh(a, step, 3, c, false, out);
}

// probes of one cell after its step
template <typename T>
__device__ FDTD_COLD void dft_probe(const StepArgs<T>& a0, long step, long c, int i, int j, int k, const T e[3],
                                    const T h[3]) {
  const StepArgs<T>& a = (const StepArgs<T>&)cold_args(a0);
  for (int p = 0; p < a.nprobe; p++)
    if (a.probe[p].cell == c) {
      const int comp = a.probe[p].comp;
      const T v = a.probe[p].field == 1 ? pick3(e[0], e[1], e[2], comp) : pick3(h[0], h[1], h[2], comp);
      a.trace[step * a.nprobe + p] = (double)v;
    }
}

// ---------------------------------------------------------------- running DFT
// F_c(w) += f_c(t_n) (cos, -sin)(w t_n) dt, E at t = (n+1) dt and H at
// (n+1.5) dt (the reference's DataCapture, include/DataCapture.hpp:61-191, is
// empty; this is the build's definition).  One thread per (monitor,
// frequency, cell) after the step, so surface work spreads over the whole GPU
// instead of serialising in the few threads that own monitor cells; it reads
// the new E and H of the monitor cells only.
// field component (0..5) recorded in slot `slot` of a monitor with component mask `mask`
__device__ __forceinline__ int comp_of_slot(int mask, int slot) {
  for (int comp = 0; comp < 6; comp++)
    if ((mask >> comp) & 1) {
      if (slot == 0) return comp;
      slot--;
    }
  return 5;
}

template <typename T>
__device__ __forceinline__ void dft_item(const StepArgs<T>& a, long g, long step) {
  int q = 0;
  while (q + 1 < a.ndft && g >= a.dft[q + 1].base) q++;
  const DftDev<T>& d = a.dft[q];
  if (step >= d.nphase) return;
  const long r = g - d.base;
  const int f = (int)(r / d.ncells);
  const long lq = r - (long)f * d.ncells;
  const int wx = d.r[1] - d.r[0], wy = d.r[3] - d.r[2];
  const int i = d.r[0] + (int)(lq % wx);
  const int j = d.r[2] + (int)((lq / wx) % wy);
  const int k = d.r[4] + (int)(lq / ((long)wx * wy));
  const long c = (long)i + a.sy * (long)j + a.sz * (long)k;
  const FDTD_G double* pe = d.phE + 2L * (d.nfreq * step + f);
  const FDTD_G double* ph = d.phH + 2L * (d.nfreq * step + f);
  const double ce = pe[0], se = pe[1], chh = ph[0], sh = ph[1];
  int slot = 0;
#pragma unroll
  for (int comp = 0; comp < 6; comp++) {
    if (!((d.mask >> comp) & 1)) continue;
    const FDTD_G T* src = comp < 3 ? a.Eo[comp] : a.Ho[comp - 3];
    const double v = src ? (double)src[c] : 0.0;
    FDTD_G double* re = d.acc + ((long)slot * d.nfreq + f) * 2 * d.ncells + lq;
    re[0] += v * (comp < 3 ? ce : chh);
    re[d.ncells] += v * (comp < 3 ? se : sh);
    slot++;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void dft_accumulate(const StepArgs<T>* __restrict__ ap, int s) {
  const StepArgs<T>& a = *ap;
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < a.dft_work) dft_item(a, g, a.step0 + s);
}

// Time-blocked form of the same sums.  dft_sample stores, after step n, the
// monitored components of every monitor cell in slot n % kDftBlock;
// dft_fold adds the samples of steps [m0, m1) (at most kDftBlock of them) to
// the accumulators, one thread per (monitor, frequency, cell) keeping its
// accumulators in registers across the block.  The additions happen in the
// same order as step-by-step accumulation, so the sums are bitwise equal.
// Each monitor's cells are padded to whole 256-thread blocks in the launch,
// so a block's monitor (and its fields) are block-uniform scalar loads.
template <typename T>
__device__ __forceinline__ void dft_sample_block(const StepArgs<T>& a, long blk, int s) {
  const long b0 = blk * 256;
  if (b0 >= a.dft_cells) return;
  int q = 0;
  while (q + 1 < a.ndft && b0 >= a.dft[q + 1].cbase) q++;
  const DftDev<T>& d = a.dft[q];
  const int lq = (int)(b0 - d.cbase) + (int)threadIdx.x;  // fewer than 2^31 cells (host-checked): 32-bit index math
  if (lq >= d.ncells) return;
  const int wx = d.r[1] - d.r[0], wy = d.r[3] - d.r[2];
  const int lj = lq / wx, i = d.r[0] + (lq - lj * wx);
  const int lk = lj / wy, j = d.r[2] + (lj - lk * wy);
  const int k = d.r[4] + lk;
  const long c = (long)i + a.sy * (long)j + a.sz * (long)k;
  const long slot0 = (long)((int)((a.step0 + s) % kDftBlock) * d.ncomp);
  int slot = 0;
#pragma unroll
  for (int comp = 0; comp < 6; comp++) {
    if (!((d.mask >> comp) & 1)) continue;
    const FDTD_G T* src = comp < 3 ? a.Eo[comp] : a.Ho[comp - 3];
    d.smp[(slot0 + slot) * d.ncells + lq] = src ? src[c] : T(0);
    slot++;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void dft_sample(const StepArgs<T>* __restrict__ ap, int s) {
  dft_sample_block(*ap, blockIdx.x, s);
}

// The fold: one thread per (component, monitor cell), cells fastest; each
// (monitor, component) segment is padded to whole 256-thread blocks, so a
// block stages the block's phase rows of its component's table in LDS once
// and every lane reads them from there.  Each sample is read once and folded
// into every frequency of its accumulators, four frequencies at a time
// (independent chains; each accumulator takes the samples in step order).
constexpr int kFoldBlock = 256;
constexpr int kFoldLdsFreq = 64;  // phase rows staged in LDS up to this many frequencies
__host__ __device__ inline long fold_segment(long ncells) { return (ncells + kFoldBlock - 1) / kFoldBlock * kFoldBlock; }

template <typename T>
__device__ __forceinline__ void dft_fold_items(const StepArgs<T>& a, long m0, long m1, long blk) {
  __shared__ double sph[kDftBlock * kFoldLdsFreq * 2];
  const long b0 = blk * kFoldBlock;  // block-uniform: one (monitor, component) per block
  if (b0 >= a.dft_slots) return;
  int q = 0;
  while (q + 1 < a.ndft && b0 >= a.dft[q + 1].sbase) q++;
  const DftDev<T>& d = a.dft[q];
  const long seg = fold_segment(d.ncells);
  const int slot = (int)((b0 - d.sbase) / seg);
  const int lq = (int)(b0 - d.sbase - (long)slot * seg) + (int)threadIdx.x;
  const int nc = (int)d.ncells;
  const long mend = m1 < (long)d.nphase ? m1 : (long)d.nphase;
  int nm = mend > m0 ? (int)(mend - m0) : 0;
  if (nm > kDftBlock) nm = kDftBlock;
  const int r0 = (int)(m0 % kDftBlock);
  const int nf = d.nfreq;
  const bool efield = comp_of_slot(d.mask, slot) < 3;
  const FDTD_G double* ph = (efield ? d.phE : d.phH) + 2L * nf * m0;
  const bool lds = nf <= kFoldLdsFreq;
  if (lds) {
    for (int x = threadIdx.x; x < nm * nf * 2; x += kFoldBlock) sph[x] = ph[x];
    __syncthreads();
  }
  if (lq >= nc) return;
  T v[kDftBlock];
#pragma unroll
  for (int t = 0; t < kDftBlock; t++)
    v[t] = t < nm ? d.smp[(long)(((r0 + t) % kDftBlock) * d.ncomp + slot) * nc + lq] : T(0);
  FDTD_G double* acc = d.acc + (long)slot * nf * 2 * nc + lq;
  for (int f0 = 0; f0 < nf; f0 += 4) {
    const int nu = nf - f0 < 4 ? nf - f0 : 4;
    double ar[4], ai[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (u < nu) {
        ar[u] = acc[(long)(f0 + u) * 2 * nc];
        ai[u] = acc[(long)(f0 + u) * 2 * nc + nc];
      }
#pragma unroll
    for (int t = 0; t < kDftBlock; t++) {
      if (t < nm) {
#pragma unroll
        for (int u = 0; u < 4; u++)
          if (u < nu) {
            const int x = (t * nf + f0 + u) * 2;
            const double pr = lds ? sph[x] : ph[x], pi = lds ? sph[x + 1] : ph[x + 1];
            ar[u] += (double)v[t] * pr;
            ai[u] += (double)v[t] * pi;
          }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (u < nu) {
        acc[(long)(f0 + u) * 2 * nc] = ar[u];
        acc[(long)(f0 + u) * 2 * nc + nc] = ai[u];
      }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void dft_fold(const StepArgs<T>* __restrict__ ap, long m0, long m1) {
  dft_fold_items(*ap, m0, m1, blockIdx.x);
}

// the fold after step s of a chunk whose step is the last of a block: the
// range comes from the staged arguments, so a captured chunk can be replayed
template <typename T>
__global__ __launch_bounds__(256) void dft_fold_step(const StepArgs<T>* __restrict__ ap, int s) {
  const StepArgs<T>& a = *ap;
  const long m1 = a.step0 + s + 1;
  const long m0 = m1 - kDftBlock > a.dft_done0 ? m1 - kDftBlock : a.dft_done0;
  dft_fold_items(a, m0, m1, blockIdx.x);
}

// the grid of batch item (or block) g: `offs` holds each grid's first one
// (prefix sums, n+1 entries)
__device__ __forceinline__ int batch_grid(const long* __restrict__ offs, int n, long g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (offs[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// the time-blocked DFT of batched grids stepped in the same phase of their
// blocks: the sampler after every step and the fold after a block's last
// step, over every grid's blocks (`boffs`: each grid's first block)
template <typename T>
__global__ __launch_bounds__(256) void dft_sample_batch(const StepArgs<T>* __restrict__ args,
                                                        const long* __restrict__ boffs, int n, int s) {
  const int q = batch_grid(boffs, n, blockIdx.x);
  dft_sample_block(args[q], blockIdx.x - boffs[q], s);
}
template <typename T>
__global__ __launch_bounds__(256) void dft_fold_batch(const StepArgs<T>* __restrict__ args,
                                                      const long* __restrict__ boffs, int n, int s) {
  const int q = batch_grid(boffs, n, blockIdx.x);
  const StepArgs<T>& a = args[q];
  const long m1 = a.step0 + s + 1;
  const long m0 = m1 - kDftBlock > a.dft_done0 ? m1 - kDftBlock : a.dft_done0;
  dft_fold_items(a, m0, m1, blockIdx.x - boffs[q]);
}

template <typename T>
__global__ __launch_bounds__(256) void dft_accumulate_batch(const StepArgs<T>* __restrict__ args,
                                                            const long* __restrict__ offs, int n, int s) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= offs[n]) return;
  const int lo = batch_grid(offs, n, g);
  const StepArgs<T>& a = args[lo];
  dft_item(a, g - offs[lo], a.step0 + s);
}

// E increment of a design cell at `step`, kept for the adjoint's eps gradient
template <typename T, bool COLD = true>
__device__ FDTD_COLD void record_increment(const StepArgs<T>& a0, long step, int i, int j, int k, T d0, T d1, T d2) {
  const StepArgs<T>& a = COLD ? (const StepArgs<T>&)cold_args(a0) : a0;
  const long row = step - a.hist_step0;
  if (row < 0 || row >= a.hist_rows) return;
  const long l = design_index(a, i, j, k);
  if (l < 0) return;
  const long nd = (long)(a.dsg[1] - a.dsg[0]) * (a.dsg[3] - a.dsg[2]) * (a.dsg[5] - a.dsg[4]);
  FDTD_G T* r = a.hist + row * 3 * nd + l;
  r[0] = d0;
  r[nd] = d1;
  r[2 * nd] = d2;
}

// Workgroup -> tile, XCD-aware.  The dispatcher deals consecutive workgroups
// round-robin to the 8 XCDs, each with its own L2.  This bijection on
// [0, nb) gives each XCD runs of kXcdRun consecutive tiles (neighbouring rows
// of one z chunk, whose shared halo rows then meet in one L2), the runs dealt
// round-robin so that unevenly loaded regions (dispersive cells, TFSF faces)
// still spread over all XCDs.  Tiles past the last whole group of 8 runs map
// to themselves.  The reverse form runs the same map from the last workgroup
// down, so a pass launched after another starts on the tiles the other ended
// with (still in the Infinity Cache).
constexpr int kXcdRun = 16;

__device__ __forceinline__ int xcd_tile(int bid, int nb) {
  const int group = 8 * kXcdRun;
  if (bid >= (nb / group) * group) return bid;
  const int x = bid % 8, loc = bid / 8;
  return ((loc / kXcdRun) * 8 + x) * kXcdRun + loc % kXcdRun;
}
__device__ __forceinline__ int xcd_tile_rev(int bid, int nb) { return xcd_tile(nb - 1 - bid, nb); }

// The 256-entry material inverses, staged in LDS once per workgroup: the
// per-cell lookup by material id then costs an LDS read instead of a global
// load that, issued behind a batch of field loads, would wait for all of them.
template <typename T>
struct MatLds {
  T ie[256], im[256];
};

template <typename T>
__device__ __forceinline__ void stage_materials(const StepArgs<T>& a, MatLds<T>& m) {
  const int n = blockDim.x * blockDim.y, t = threadIdx.x + blockDim.x * threadIdx.y;
  for (int q = t; q < 256; q += n) {
    m.ie[q] = a.mie[q];
    m.im[q] = a.mim[q];
  }
  __syncthreads();
}

// ---------------------------------------------------------------- boundary conditions
// BoundaryUpdates.hpp: SingleUpdate::atomic_update (258-300) sets, for every
// component of one field type, ghost = Mask * source, the ghost being the
// boundary cell's outward neighbour; the source is the boundary cell itself
// for PEC / PMC / Symmetric / Antisymmetric (masks 527-722) or its periodic
// image (800-832, mask 1).  Placement (which the reference leaves to the
// caller): ghost layer 0 / L-1, boundary layer 1 / L-2, periodic image of
// ghost 0 layer L-2 and of ghost L-1 layer 1; tangential range the interior
// cells.  The pass that computes a source cell's new value also writes the
// ghost copies (masks are +-1, so they are exact): nothing in the same pass
// reads a ghost of the field it writes, and the next pass sees them.
// ghost copies of interior cell (i,j,k) (address c) whose new values are v
// (ft 0: E, 1: H); the components of the mode (non-null field pointers)
template <typename T, bool COLD = true>
__device__ FDTD_COLD void bc_ghost_write(const StepArgs<T>& a0, int ft, int ndim, int i, int j, int k, long c,
                                         const T v[3]) {
  const StepArgs<T>& a = COLD ? (const StepArgs<T>&)cold_args(a0) : a0;
  const int ci[3] = {i, j, k};
  FDTD_G T* const* F = ft == 0 ? a.E : a.H;  // callers are the in-place passes
  for (int ax = 0; ax < ndim; ax++)
    for (int side = 0; side < 2; side++) {
      if (ci[ax] != a.bc_src[ax][side]) continue;
      const long d = c + a.bc_off[ax][side];
      const int neg = a.bc_neg[ft][ax][side];
      for (int q = 0; q < 3; q++) {
        if (!F[q]) continue;  // a component the mode does not have
        F[q][d] = ((neg >> q) & 1) ? -v[q] : v[q];
      }
    }
}

// ---------------------------------------------------------------- two passes
// One thread per cell and z plane: every load of the cell is issued before
// any arithmetic, so a wave carries ~12 independent loads in flight.
// GEN: the general instance, for grids with boundary conditions (ghost
// writes), E / H soft sources or single TFSF planes; DISP: for grids with
// dispersive materials (1: ADE poles, 2: also magnetized and fluid Drude).
// An instance without them carries none of that code, so its registers stay
// low.
template <typename T, bool GEN = false, int DISP = 1>
__device__ __forceinline__ void e_pass_cells(const StepArgs<T>& a, const T* tie, int i, int j, int k0, int k1, int s) {
  if (i < 1 || i >= a.nx - 1 || j >= a.ny - 1) return;
  const FDTD_G T* __restrict__ Hx = a.H[0];
  const FDTD_G T* __restrict__ Hy = a.H[1];
  const FDTD_G T* __restrict__ Hz = a.H[2];
  const FDTD_G T* __restrict__ Ex = a.E[0];
  const FDTD_G T* __restrict__ Ey = a.E[1];
  const FDTD_G T* __restrict__ Ez = a.E[2];
  const long sy = a.sy, sz = a.sz;
  const long step = a.step0 + s;
  const FDTD_G double* incH = a.incH + (long)s * a.auxL;
  const T fx = a.cf[0][0][i], fy = a.cf[0][1][j];
  // CPML inputs, gathered with the field loads so that their loads overlap
  // them (in the cell update they would follow the curl arithmetic): the
  // slab tests are per thread along x and y, uniform along z.  FP32 only:
  // the FP64 passes would lose a wave per SIMD to the registers (measured
  // 4% faster at 256^3 FP32)
  constexpr int PF = sizeof(T) == 4 ? 1 : 0;
  CellPre<T> pre;
  pre.src = nullptr;
  pre.nsrc = 0;
  const int sxi = PF ? slab_index(i, a.slab_a[0], a.slab_b[0]) : -1;
  const int syi = PF ? slab_index(j, a.slab_a[1], a.slab_b[1]) : -1;
  if (sxi >= 0) {
    pre.g[0] = a.cg[0][0][i];
    pre.b[0] = a.cb[0][0][i];
    pre.c[0] = a.cc[0][0][i];
  }
  if (syi >= 0) {
    pre.g[1] = a.cg[0][1][j];
    pre.b[1] = a.cb[0][1][j];
    pre.c[1] = a.cc[0][1][j];
  }
  // the old H of the plane below (the z-1 neighbours of the E update): the
  // march carries the cell's Hx, Hy from plane to plane instead of reloading
  T hx_zm = Hx[(long)i + sy * (long)j + sz * (long)(k0 - 1)], hy_zm = Hy[(long)i + sy * (long)j + sz * (long)(k0 - 1)];
  for (int k = k0; k < k1; k++) {
    const long c = (long)i + sy * (long)j + sz * (long)k;
    if (sxi >= 0) {
      const long o = slab_off(a, 0, sxi, i, j, k);
      pre.psi[0][0] = a.psi[0][1][0][o];
      pre.psi[0][1] = a.psi[0][2][0][o];
    }
    if (syi >= 0) {
      const long o = slab_off(a, 1, syi, i, j, k);
      pre.psi[1][0] = a.psi[0][0][1][o];
      pre.psi[1][1] = a.psi[0][2][1][o];
    }
    const int szi = PF ? slab_index(k, a.slab_a[2], a.slab_b[2]) : -1;
    if (szi >= 0) {
      const long o = slab_off(a, 2, szi, i, j, k);
      pre.psi[2][0] = a.psi[0][0][2][o];
      pre.psi[2][1] = a.psi[0][1][2][o];
      pre.g[2] = a.cg[0][2][k];
      pre.b[2] = a.cb[0][2][k];
      pre.c[2] = a.cc[0][2][k];
    }
    HStencil<T> h;
    h.hx = Hx[c];
    h.hy = Hy[c];
    h.hz = Hz[c];
    h.hz_ym = Hz[c - sy];
    h.hz_xm = Hz[c - 1];
    h.hy_xm = Hy[c - 1];
    // the general instance reloads the plane below (a cache hit) instead: its
    // carried pair, loaded before the loop, was copied at the loop header,
    // and the wait for those loads there also waited, every plane, for the
    // previous plane's stores (measured 10 % slower on a periodic 256^3 grid)
    h.hy_zm = GEN ? Hy[c - sz] : hy_zm;
    h.hx_ym = Hx[c - sy];
    h.hx_zm = GEN ? Hx[c - sz] : hx_zm;
    hx_zm = h.hx;
    hy_zm = h.hy;
    const T ex = Ex[c], ey = Ey[c], ez = Ez[c];
    const int mid = a.mat[c];
    const T fz = a.cf[0][2][k];
    T o[3];
    e_cell_3d<T, DISP, true, PF, const FDTD_G double*, GEN>(a, step, incH, i, j, k, c, h, ex, ey, ez, mid,
                                                             cell_inv_eps(a, tie, mid, i, j, k), fx, fy, fz, true, o,
                                                             &pre);
    a.Eo[0][c] = o[0];
    a.Eo[1][c] = o[1];
    a.Eo[2][c] = o[2];
    if (a.hist) record_increment<T, true>(a, step, i, j, k, o[0] - ex, o[1] - ey, o[2] - ez);
    if (GEN && a.bc_any) bc_ghost_write<T, true>(a, 0, 3, i, j, k, c, o);
  }
}

template <typename T, typename Tab, bool GEN = false>
__device__ __forceinline__ void h_pass_cells(const StepArgs<T>& a, Tab tim, int i, int j, int k0, int k1, int s) {
  if (i < 1 || i >= a.nx - 1 || j >= a.ny - 1) return;
  // in place: the E pass wrote E; H is read and written through the same pointers
  const FDTD_G T* __restrict__ Ex = a.E[0];
  const FDTD_G T* __restrict__ Ey = a.E[1];
  const FDTD_G T* __restrict__ Ez = a.E[2];
  FDTD_G T* __restrict__ Hx = a.H[0];
  FDTD_G T* __restrict__ Hy = a.H[1];
  FDTD_G T* __restrict__ Hz = a.H[2];
  const long sy = a.sy, sz = a.sz;
  const long step = a.step0 + s;
  const FDTD_G double* incE = a.incE + (long)s * a.auxL;
  const bool extra = a.nprobe > 0;
  const T fx = a.cf[1][0][i], fy = a.cf[1][1][j];
  // CPML inputs gathered with the field loads, as in the E pass (FP32)
  constexpr int PF = sizeof(T) == 4 ? 1 : 0;
  CellPre<T> pre;
  pre.src = nullptr;
  pre.nsrc = 0;
  const int sxi = PF ? slab_index(i, a.slab_a[0], a.slab_b[0]) : -1;
  const int syi = PF ? slab_index(j, a.slab_a[1], a.slab_b[1]) : -1;
  if (sxi >= 0) {
    pre.g[0] = a.cg[1][0][i];
    pre.b[0] = a.cb[1][0][i];
    pre.c[0] = a.cc[1][0][i];
  }
  if (syi >= 0) {
    pre.g[1] = a.cg[1][1][j];
    pre.b[1] = a.cb[1][1][j];
    pre.c[1] = a.cc[1][1][j];
  }
  for (int k = k0; k < k1; k++) {
    const long c = (long)i + sy * (long)j + sz * (long)k;
    if (sxi >= 0) {
      const long o = slab_off(a, 0, sxi, i, j, k);
      pre.psi[0][0] = a.psi[1][1][0][o];
      pre.psi[0][1] = a.psi[1][2][0][o];
    }
    if (syi >= 0) {
      const long o = slab_off(a, 1, syi, i, j, k);
      pre.psi[1][0] = a.psi[1][0][1][o];
      pre.psi[1][1] = a.psi[1][2][1][o];
    }
    const int szi = PF ? slab_index(k, a.slab_a[2], a.slab_b[2]) : -1;
    if (szi >= 0) {
      const long o = slab_off(a, 2, szi, i, j, k);
      pre.psi[2][0] = a.psi[1][0][2][o];
      pre.psi[2][1] = a.psi[1][1][2][o];
      pre.g[2] = a.cg[1][2][k];
      pre.b[2] = a.cb[1][2][k];
      pre.c[2] = a.cc[1][2][k];
    }
    EStencil<T> e;
    e.ex = Ex[c];
    e.ey = Ey[c];
    e.ez = Ez[c];
    e.ez_yp = Ez[c + sy];
    e.ez_xp = Ez[c + 1];
    e.ey_xp = Ey[c + 1];
    e.ey_zp = Ey[c + sz];
    e.ex_yp = Ex[c + sy];
    e.ex_zp = Ex[c + sz];
    const T hx = Hx[c], hy = Hy[c], hz = Hz[c];
    const int mid = a.mat[c];
    const T fz = a.cf[1][2][k];
    T o[3];
    h_cell_3d<T, true, PF, const FDTD_G double*, GEN>(a, step, incE, i, j, k, c, e, hx, hy, hz, tim[mid], fx, fy, fz,
                                                       o, &pre);
    Hx[c] = o[0];
    Hy[c] = o[1];
    Hz[c] = o[2];
    if (GEN && a.bc_any) bc_ghost_write(a, 1, 3, i, j, k, c, o);
    if (extra) {
      const T ev[3] = {e.ex, e.ey, e.ez};
      dft_probe(a, step, c, i, j, k, ev, o);
    }
  }
}

// One-dimensional launch of ntx x nty x nkc tiles of blockDim.x x blockDim.y
// columns and kc planes, XCD-aware (tiles x fastest, then y, then z chunks).
template <typename T, bool GEN = false, int DISP = 1>
__global__ __launch_bounds__(256) void e_pass_3d(const StepArgs<T>* __restrict__ ap, int kc, int s, int ntx, int nty,
                                                 int nkc) {
  const StepArgs<T>& a = *ap;
  __shared__ MatLds<T> sm;
  stage_materials(a, sm);
  const int t = xcd_tile(blockIdx.x, ntx * nty * nkc);
  const int bx = t % ntx, by = (t / ntx) % nty, bz = t / (ntx * nty);
  const int k0 = 1 + bz * kc;
  e_pass_cells<T, GEN, DISP>(a, sm.ie, bx * blockDim.x + threadIdx.x, 1 + by * blockDim.y + threadIdx.y, k0,
                            min(k0 + kc, a.nz - 1), s);
}

// Each XCD walks its tile range in reverse: its first tiles read the E and H
// that its E pass touched last (still in its L2 / the Infinity Cache); the
// next E pass then starts where this pass ended.
//
// FP32 instances are compiled for 7 waves per SIMD (the register count they
// need anyway): at 256^3 on MI355X that schedule runs the H pass ~10 % faster
// (0.151 vs 0.168 ms per step) than the compiler's default.
template <typename T, bool GEN>
__device__ __forceinline__ void h_pass_3d_body(const StepArgs<T>* __restrict__ ap, int kc, int s, int ntx, int nty,
                                               int nkc) {
  const StepArgs<T>& a = *ap;
  const int t = xcd_tile_rev(blockIdx.x, ntx * nty * nkc);
  const int bx = t % ntx, by = (t / ntx) % nty, bz = t / (ntx * nty);
  const int k0 = 1 + bz * kc;
  // the H pass reads 1/(mu_r mu0) from global memory: staging it in LDS
  // measured slower here (the E pass, with its larger stencil, gains)
  h_pass_cells<T, decltype(a.mim), GEN>(a, a.mim, bx * blockDim.x + threadIdx.x, 1 + by * blockDim.y + threadIdx.y,
                                       k0, min(k0 + kc, a.nz - 1), s);
}

template <typename T, bool GEN = false>
__global__ __launch_bounds__(256) void h_pass_3d(const StepArgs<T>* __restrict__ ap, int kc, int s, int ntx, int nty,
                                                 int nkc) {
  h_pass_3d_body<T, GEN>(ap, kc, s, ntx, nty, nkc);
}
template <>
__global__ __attribute__((amdgpu_waves_per_eu(7))) __launch_bounds__(256) void h_pass_3d<float, false>(
    const StepArgs<float>* __restrict__ ap, int kc, int s, int ntx, int nty, int nkc) {
  h_pass_3d_body<float, false>(ap, kc, s, ntx, nty, nkc);
}
// (the general instance at 6 waves: at 7 it spills 68 bytes and runs 4 % slower)
template <>
__global__ __attribute__((amdgpu_waves_per_eu(6))) __launch_bounds__(256) void h_pass_3d<float, true>(
    const StepArgs<float>* __restrict__ ap, int kc, int s, int ntx, int nty, int nkc) {
  h_pass_3d_body<float, true>(ap, kc, s, ntx, nty, nkc);
}

// ---------------------------------------------------------------- Bloch-periodic faces
// A complex field is held as two real parts, each stepped by the passes with
// its own arguments (re, im).  Every Bloch face writes its ghost layer from
// the periodic image of both parts: ghost = f * image with f = fr + i fi for
// E and conj(f) for H (BlochDecorator and make_bloch_periodic_boundary,
// BoundaryUpdates.hpp:845-885; the mask of this kind is 1 for every
// component).  One thread per tangential interior cell of a face (blockIdx.y
// = face); run after both parts' E pass (ft 0) and after both H passes (ft 1).
struct BlochFace {
  int axis, src, dst;  // normal axis, image layer, ghost layer
  double fr, fi;       // phase factor of E (H takes its conjugate)
};
struct BlochFaces {
  int n;
  BlochFace f[6];
  int len[3];
};
template <typename T>
__global__ __launch_bounds__(256) void bloch_ghosts(const StepArgs<T>* __restrict__ re,
                                                    const StepArgs<T>* __restrict__ im, BlochFaces bf, int ft) {
  const BlochFace fc = bf.f[blockIdx.y];
  const int a = fc.axis, t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
  const int n1 = bf.len[t1] - 2, n2 = bf.len[t2] > 1 ? bf.len[t2] - 2 : 1;  // TM: one plane, k = 0
  const long u = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u >= long(n1) * n2) return;
  int c[3];
  c[t1] = 1 + int(u % n1);
  c[t2] = bf.len[t2] > 1 ? 1 + int(u / n1) : 0;
  c[a] = fc.src;
  const long sy = bf.len[0], sz = long(bf.len[0]) * bf.len[1];
  const long q = c[0] + sy * c[1] + sz * c[2];
  c[a] = fc.dst;
  const long d = c[0] + sy * c[1] + sz * c[2];
  const T fr = T(fc.fr), fi = ft == 0 ? T(fc.fi) : T(-fc.fi);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    FDTD_G T* xr = ft == 0 ? re->E[k] : re->H[k];
    FDTD_G T* xi = ft == 0 ? im->E[k] : im->H[k];
    if (!xr) continue;
    const T sr = xr[q], si = xi[q];
    xr[d] = fr * sr - fi * si;
    xi[d] = fi * sr + fr * si;
  }
}

// ---------------------------------------------------------------- surface quantities
// UpdateSurfaceChargeDensity (SurfaceQuantities.hpp:174-183), UpdateSurfaceCurrentE
// (233-244; TM 258-266) and UpdateSurfaceCurrentH (293-304; TM 319-327) on the
// current E and H of the cells of region [x0,x1)x[y0,y1)x[z0,z1): one thread
// per cell, FP64 arithmetic in the reference's operation order with no FMA
// contraction, so FP64 grids give the reference's bits.  out: quantity-major,
// x fastest; 3-D: rho, Jex, Jey, Jez, Jhx, Jhy, Jhz; TM: Jez, Jhx, Jhy; TE
// (189-198, 253-263, 323-332): rho, Jex, Jey, Jhz.  mode: 3, 2 (TM) or 4 (TE).
struct SurfRegion {
  int x0, x1, y0, y1, z0, z1;
};

template <typename T>
__global__ __launch_bounds__(256) void surface_quantities(const T* __restrict__ Ex, const T* __restrict__ Ey,
                                                          const T* __restrict__ Ez, const T* __restrict__ Hx,
                                                          const T* __restrict__ Hy, const T* __restrict__ Hz, long sy,
                                                          long sz, SurfRegion r, int mode, double dx, double eps0,
                                                          double* __restrict__ out) {
#pragma clang fp contract(off)
  const long ex = r.x1 - r.x0, ey = r.y1 - r.y0, ez = r.z1 - r.z0;
  const long cells = ex * ey * ez;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= cells) return;
  const int i = r.x0 + int(q % ex), j = r.y0 + int((q / ex) % ey), k = r.z0 + int(q / (ex * ey));
  const long c = (long)i + sy * (long)j + sz * (long)k;
  auto v = [](const T* f, long x) { return (double)f[x]; };
  if (mode == 3) {
    out[0 * cells + q] = eps0 / dx *
                         (v(Ex, c) + v(Ey, c) + v(Ez, c) - v(Ex, c - 1) - v(Ey, c - sy) - v(Ez, c - sz));
    out[1 * cells + q] = 1.0 / dx * (v(Hz, c) - v(Hz, c - sy) - v(Hy, c) + v(Hy, c - sz));
    out[2 * cells + q] = 1.0 / dx * (v(Hx, c) - v(Hx, c - sz) - v(Hz, c) + v(Hz, c - 1));
    out[3 * cells + q] = 1.0 / dx * (v(Hy, c) - v(Hy, c - 1) - v(Hx, c) + v(Hx, c - sy));
    out[4 * cells + q] = -1.0 / dx * (v(Ez, c + sy) - v(Ez, c) - v(Ey, c + sz) + v(Ey, c));
    out[5 * cells + q] = -1.0 / dx * (v(Ex, c + sz) - v(Ex, c) - v(Ez, c + 1) + v(Ez, c));
    out[6 * cells + q] = -1.0 / dx * (v(Ey, c + 1) - v(Ey, c) - v(Ex, c + sy) + v(Ex, c));
  } else if (mode == 2) {
    out[0 * cells + q] = 1.0 / dx * (v(Hy, c) - v(Hy, c - 1) - v(Hx, c) + v(Hx, c - sy));
    out[1 * cells + q] = -1.0 / dx * (v(Ez, c + sy) - v(Ez, c));
    out[2 * cells + q] = -1.0 / dx * (-v(Ez, c + 1) + v(Ez, c));
  } else {
    out[0 * cells + q] = eps0 / dx * (v(Ex, c) + v(Ey, c) - v(Ex, c - 1) - v(Ey, c - sy));
    out[1 * cells + q] = 1.0 / dx * (v(Hz, c) - v(Hz, c - sy));
    out[2 * cells + q] = 1.0 / dx * (-v(Hz, c) + v(Hz, c - 1));
    out[3 * cells + q] = -1.0 / dx * (v(Ey, c + 1) - v(Ey, c) - v(Ex, c + sy) + v(Ex, c));
  }
}

// ---------------------------------------------------------------- batched passes
// Many independent grids in one launch per pass: workgroup b takes an entry
// (grid index, tile x, tile y, tile z) of a table built on the host, in the
// XCD-aware order of the single-grid passes, each grid reading its own
// arguments.  Same cell arithmetic as above.
// DISP: 0 when no grid of the batch holds a dispersive material (no ADE code)
template <typename T, int DISP = 1>
__global__ __launch_bounds__(256) void e_pass_batch(const StepArgs<T>* __restrict__ args, const int4* __restrict__ tiles,
                                                    int kc, int s) {
  const int4 t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const StepArgs<T>& a = args[t.x];
  __shared__ MatLds<T> sm;
  stage_materials(a, sm);
  const int k0 = 1 + t.w * kc;
  e_pass_cells<T, false, DISP>(a, sm.ie, t.y * blockDim.x + threadIdx.x, 1 + t.z * blockDim.y + threadIdx.y, k0,
                               min(k0 + kc, a.nz - 1), s);
}

template <typename T>
__device__ __forceinline__ void h_pass_batch_body(const StepArgs<T>* __restrict__ args, const int4* __restrict__ tiles,
                                                  int kc, int s) {
  const int4 t = tiles[xcd_tile_rev(blockIdx.x, gridDim.x)];
  const StepArgs<T>& a = args[t.x];
  const int k0 = 1 + t.w * kc;
  h_pass_cells(a, a.mim, t.y * blockDim.x + threadIdx.x, 1 + t.z * blockDim.y + threadIdx.y, k0, min(k0 + kc, a.nz - 1),
               s);
}
template <typename T>
__global__ __launch_bounds__(256) void h_pass_batch(const StepArgs<T>* __restrict__ args, const int4* __restrict__ tiles,
                                                    int kc, int s) {
  h_pass_batch_body(args, tiles, kc, s);
}
// FP32: the single-grid H pass's 7-waves-per-SIMD schedule
template <>
__global__ __attribute__((amdgpu_waves_per_eu(7))) __launch_bounds__(256) void h_pass_batch<float>(
    const StepArgs<float>* __restrict__ args, const int4* __restrict__ tiles, int kc, int s) {
  h_pass_batch_body(args, tiles, kc, s);
}

// Persistent form of the batched passes: gridDim.x workgroups (a multiple of
// 8) walk the tile table, workgroup b taking the tiles the one-per-tile launch
// gives blocks b, b + gridDim.x, b + 2 gridDim.x, ... (so each XCD keeps its
// tiles and its neighbours' timing); the E pass stages a grid's material
// table in LDS only when its next tile belongs to another grid.  Same cell
// code and order per cell: bitwise the one-per-tile result.
template <typename T, int DISP = 1>
__global__ __launch_bounds__(256) void e_pass_batch_p(const StepArgs<T>* __restrict__ args,
                                                      const int4* __restrict__ tiles, int ntiles, int kc, int s) {
  __shared__ MatLds<T> sm;
  int cur = -1;
  for (int v = blockIdx.x; v < ntiles; v += gridDim.x) {
    const int4 t = tiles[xcd_tile(v, ntiles)];
    const StepArgs<T>& a = args[t.x];
    if (t.x != cur) {  // uniform over the workgroup
      if (cur >= 0) __syncthreads();  // every wave is done with the previous table
      stage_materials(a, sm);
      cur = t.x;
    }
    const int k0 = 1 + t.w * kc;
    e_pass_cells<T, false, DISP>(a, sm.ie, t.y * blockDim.x + threadIdx.x, 1 + t.z * blockDim.y + threadIdx.y, k0,
                                 min(k0 + kc, a.nz - 1), s);
  }
}
template <typename T>
__device__ __forceinline__ void h_pass_batch_p_body(const StepArgs<T>* __restrict__ args,
                                                    const int4* __restrict__ tiles, int ntiles, int kc, int s) {
  for (int v = blockIdx.x; v < ntiles; v += gridDim.x) {
    const int4 t = tiles[xcd_tile_rev(v, ntiles)];
    const StepArgs<T>& a = args[t.x];
    const int k0 = 1 + t.w * kc;
    h_pass_cells(a, a.mim, t.y * blockDim.x + threadIdx.x, 1 + t.z * blockDim.y + threadIdx.y, k0,
                 min(k0 + kc, a.nz - 1), s);
  }
}
template <typename T>
__global__ __launch_bounds__(256) void h_pass_batch_p(const StepArgs<T>* __restrict__ args,
                                                      const int4* __restrict__ tiles, int ntiles, int kc, int s) {
  h_pass_batch_p_body(args, tiles, ntiles, kc, s);
}
template <>
__global__ __attribute__((amdgpu_waves_per_eu(7))) __launch_bounds__(256) void h_pass_batch_p<float>(
    const StepArgs<float>* __restrict__ args, const int4* __restrict__ tiles, int ntiles, int kc, int s) {
  h_pass_batch_p_body(args, tiles, ntiles, kc, s);
}

// ---------------------------------------------------------------- buffer access
// Raw buffer access (32-bit byte offsets: one lane offset shared by every
// array, the row/plane part in a scalar register), which keeps the many rows
// of a wave's brick from each holding 64-bit addresses per array.  Every
// access lands inside the (padded) allocation.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const FDTD_G void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so);
template <>
__device__ __forceinline__ float buf_ld<float>(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0));
}
template <>
__device__ __forceinline__ double buf_ld<double>(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, 0));
}
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)vo, (int)so, 0);
}
__device__ __forceinline__ void buf_st(double v, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v),
                                        r, (int)vo, (int)so, 0);
}


// ---------------------------------------------------------------- streamed fused pass
// One pass per step (FP32), every E and H value read once and written once
// per step.  A tile is an E region of kSCols x kSRows cells (x, y) marched
// along z; its owned cells are the first kSCols-1 columns and kSRows-1 rows,
// the last column and row being the +x / +y halo whose new E the owned cells'
// H update needs.  Per plane k of the march: new E of plane k+1 (halo cells
// too, from the old H of planes k+1 and k), then new H of plane k from the new
// E of planes k (LDS: +x, +y neighbours) and k+1 (the thread's registers).
// Old state is only read and new state only written (ping-pong copies), so
// halo cells that two tiles recompute never race.  The cell arithmetic is
// e_cell_3d / h_cell_3d, as in the two-pass schedule, so the fields are
// bitwise equal.
//
// The waves of a workgroup have two roles.  Two producer waves stream the old
// H of each plane, with the -x / -y halo the E update reads, into a ring of
// LDS planes by direct buffer-to-LDS loads, kSDepth planes ahead of the plane
// the consumers need; they issue nothing else, so a wait for "all but the
// newest n groups" of their loads is exact.  The consumer waves (one cell per
// thread) gather their own cell's old E, material id and CPML psi a plane
// ahead into registers, compute, and store without ever waiting for their
// stores: the barriers between planes wait only for LDS operations and for
// the producers' loads.  The z coefficients of the grid and the step's
// incident rows sit in LDS, so no update waits on a scalar load.
//
// The kernel is persistent: gridDim.x workgroups split the (tile, plane)
// units evenly, each taking a contiguous range (a run of planes of one tile,
// possibly ending one tile and starting the next), so every workgroup does the
// same work and the launch has no tail.
constexpr int kSCols = 128;                 // E-region columns (two waves across); 127 owned
constexpr int kSRows = 5;                   // E-region rows; 4 owned
constexpr int kSCons = 2 * kSRows;          // consumer waves, one cell per thread
constexpr int kSProd = 2;                   // producer waves
constexpr int kSThreads = 64 * (kSCons + kSProd);
constexpr int kSDepth = 3;                  // planes of old H in flight ahead of the one needed
constexpr int kSSlots = kSDepth + 2;        // plus the two planes being read
constexpr int kSMaxAux = 1024;              // incident rows staged in LDS up to this length
constexpr int kSMaxZ = 1024;                // z coefficients staged in LDS up to this many planes

__host__ __device__ constexpr int pad64(int n) { return (n + 63) / 64 * 64; }
struct SGeom {
  static constexpr int nHx = (kSRows + 1) * kSCols;        // rows j0-1 .. j0+R-1, columns i0 .. i0+127
  static constexpr int nHy = kSRows * (kSCols + 1);        // rows j0 .. j0+R-1, columns i0-1 .. i0+127
  static constexpr int nHz = (kSRows + 1) * (kSCols + 1);  // rows j0-1 .. j0+R-1, columns i0-1 .. i0+127
  static constexpr int nId = kSRows * kSCols;              // material ids of the E region
  static constexpr int qHx = pad64(nHx) / 64, qHy = pad64(nHy) / 64, qHz = pad64(nHz) / 64, qId = pad64(nId) / 64;
  static constexpr int nInst = qHx + qHy + qHz + qId;  // 64-lane loads per plane
  static constexpr int perProd = nInst / kSProd;
  static constexpr int oHy = pad64(nHx), oHz = oHy + pad64(nHy), slot = oHz + pad64(nHz);
};
static_assert(SGeom::nInst % kSProd == 0, "producer waves issue equal shares");
static_assert(SGeom::perProd * (kSDepth - 1) <= 63, "a producer's wait count fits vmcnt");
static_assert(kSThreads <= 1024, "workgroup size");

struct StreamLds {
  float h[kSSlots][SGeom::slot];          // ring of old-H planes
  int id[kSSlots][pad64(SGeom::nId)];     // their material ids (one byte loaded into each int)
  float en[3][kSRows * kSCols];           // new E of the plane below the one being computed
  float zc[2][4][kSMaxZ];                 // per field type and plane: dt F/dx, dt G, B, C/dx
  double inc[2][kSMaxAux];                // the step's incident rows: aux H (E update), aux E (H update)
  SrcVal<float> src[kMaxSrc];             // the step's soft-source values
  ProbeDev probe[kMaxProbe];
  int nsrc;
  MatLds<float> m;
};

__device__ __forceinline__ void lds_dma(__amdgpu_buffer_rsrc_t r, void* lds, int size, unsigned vo, unsigned so) {
  if (size == 4)
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)lds, 4, (int)vo, (int)so, 0, 0);
  else
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)lds, 1, (int)vo, (int)so, 0, 0);
}
// workgroup barriers that wait only for what the wave's role needs (a
// __syncthreads fence would also wait for every store and load in flight)
__device__ __forceinline__ void s_bar_lds() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }
template <int N>
__device__ __forceinline__ void s_bar_vm() { asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"(N) : "memory"); }

// the cell data a consumer gathers ahead of its use
struct SPre {
  float eo[3];               // old E of the E plane
  float pe[3][2], ph[3][2];  // CPML psi of the E plane and of the H plane
};

template <int DISP>
__global__ __launch_bounds__(kSThreads) void fused3d_lds(const StepArgs<float>* __restrict__ ap, int s, int ntx,
                                                          long units) {
  using T = float;
  using Gm = SGeom;
  const StepArgs<T>& a = *ap;
  __shared__ StreamLds L;
  const int nx = a.nx, ny = a.ny, nz = a.nz, nzi = nz - 2;
  const long sy = a.sy, sz = a.sz;
  const long step = a.step0 + s;
  // stage the material table, the z coefficients, the step's incident rows,
  // source values and the probe table
  stage_materials(a, L.m);
  for (int q = threadIdx.x; q < 8 * nz; q += blockDim.x) {
    const int ft = q / (4 * nz), w = (q / nz) % 4, kk = q % nz;
    const FDTD_G T* src = w == 0 ? a.cf[ft][2] : (w == 1 ? a.cg[ft][2] : (w == 2 ? a.cb[ft][2] : a.cc[ft][2]));
    L.zc[ft][w][kk] = src[kk];
  }
  const bool tfsf_lds = a.tfsf && a.auxL <= kSMaxAux;
  if (tfsf_lds)
    for (int q = threadIdx.x; q < a.auxL; q += blockDim.x) {
      L.inc[0][q] = a.incH[(long)s * a.auxL + q];
      L.inc[1][q] = a.incE[(long)s * a.auxL + q];
    }
  if (threadIdx.x == 0) {
    int n = 0;
    for (int q = 0; q < a.nsrc; q++)  // add_sources adds a source while step < its table length
      if (step < a.src[q].n) L.src[n++] = SrcVal<T>{a.src[q].cell, a.src[q].field, a.src[q].comp, (T)a.src[q].table[step]};
    L.nsrc = n;
  }
  for (int q = threadIdx.x; q < a.nprobe; q += blockDim.x) L.probe[q] = a.probe[q];
  __syncthreads();
  const unsigned plane_bytes = (unsigned)(sz * (long)sizeof(T));
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const long u0 = units * (long)blockIdx.x / gridDim.x, u1 = units * ((long)blockIdx.x + 1) / gridDim.x;
  if (w >= kSCons) {
    // ------------------------------------------------ producer waves
    const int pw = w - kSCons;
    __amdgpu_buffer_rsrc_t rH[3];
#pragma unroll
    for (int q = 0; q < 3; q++) rH[q] = buf_rsrc(a.H[q], a.field_bytes);
    const __amdgpu_buffer_rsrc_t rM = buf_rsrc(a.mat, a.mat_bytes);
    for (long u = u0; u < u1;) {
      const int t = (int)(u / nzi);
      const long uend = min(u1, (long)(t + 1) * nzi);
      const int ka = 1 + (int)(u - (long)t * nzi), kb = 1 + (int)(uend - (long)t * nzi);
      u = uend;
      const int i0 = 1 + (t % ntx) * (kSCols - 1), j0 = 1 + (t / ntx) * (kSRows - 1);
      unsigned vo[Gm::perProd];  // byte offsets (plane 0) of this lane's elements, per load
#pragma unroll
      for (int q = 0; q < Gm::perProd; q++) {
        const int g = pw * Gm::perProd + q;
        int e, x, y;
        if (g < Gm::qHx) {
          e = min(g * 64 + lane, Gm::nHx - 1);
          x = i0 + e % kSCols;
          y = j0 - 1 + e / kSCols;
        } else if (g < Gm::qHx + Gm::qHy) {
          e = min((g - Gm::qHx) * 64 + lane, Gm::nHy - 1);
          x = i0 - 1 + e % (kSCols + 1);
          y = j0 + e / (kSCols + 1);
        } else if (g < Gm::qHx + Gm::qHy + Gm::qHz) {
          e = min((g - Gm::qHx - Gm::qHy) * 64 + lane, Gm::nHz - 1);
          x = i0 - 1 + e % (kSCols + 1);
          y = j0 - 1 + e / (kSCols + 1);
        } else {
          e = min((g - Gm::qHx - Gm::qHy - Gm::qHz) * 64 + lane, Gm::nId - 1);
          x = i0 + e % kSCols;
          y = j0 + e / kSCols;
        }
        const unsigned cell = (unsigned)((long)min(x, nx - 1) + sy * (long)min(y, ny - 1));
        vo[q] = g < Gm::qHx + Gm::qHy + Gm::qHz ? cell * (unsigned)sizeof(T) : cell;
      }
      auto issue = [&](int k) {  // old H and material ids of plane k into ring slot k % kSSlots
        const int sl = k % kSSlots;
#pragma unroll
        for (int q = 0; q < Gm::perProd; q++) {
          const int g = pw * Gm::perProd + q;  // uniform per wave, and so is the choice of array
          if (g < Gm::qHx)
            lds_dma(rH[0], &L.h[sl][g * 64], 4, vo[q], (unsigned)k * plane_bytes);
          else if (g < Gm::qHx + Gm::qHy)
            lds_dma(rH[1], &L.h[sl][Gm::oHy + (g - Gm::qHx) * 64], 4, vo[q], (unsigned)k * plane_bytes);
          else if (g < Gm::qHx + Gm::qHy + Gm::qHz)
            lds_dma(rH[2], &L.h[sl][Gm::oHz + (g - Gm::qHx - Gm::qHy) * 64], 4, vo[q], (unsigned)k * plane_bytes);
          else
            lds_dma(rM, &L.id[sl][(g - Gm::qHx - Gm::qHy - Gm::qHz) * 64], 1, vo[q], (unsigned)(k * sz));
        }
      };
      // planes ka-1 .. ka+3 (up to kb) fill the ring; the prologue needs ka-1 and ka
      int issued = ka - 2;  // last plane issued
      for (int p = ka - 1; p <= min(ka + 3, kb); p++) issue(issued = p);
      // barrier once every plane <= need has landed (at most the two newest
      // groups, planes need+1 and need+2, may stay in flight)
      auto wait_bar = [&](int need) {
        const int after = issued - need;
        if (after >= 2)
          s_bar_vm<Gm::perProd * 2>();
        else if (after == 1)
          s_bar_vm<Gm::perProd>();
        else
          s_bar_vm<0>();
      };
      wait_bar(ka);  // B0
      for (int k = ka; k < kb; k++) {
        wait_bar(k + 1);                         // B1: plane k+1 landed (k+2, k+3 in flight)
        if (k + 4 <= kb) issue(issued = k + 4);  // into the slot of plane k-1, free since B2 of the last plane
        s_bar_lds();                             // B2
      }
      s_bar_vm<0>();  // the range's end
    }
    return;
  }
  // ------------------------------------------------ consumer waves: one E-region cell per thread
  const int cx = 64 * (w & 1) + lane, ry = w >> 1;
  __amdgpu_buffer_rsrc_t rE[3];
#pragma unroll
  for (int q = 0; q < 3; q++) rE[q] = buf_rsrc(a.E[q], a.field_bytes);
  const int nprobe = a.nprobe;
  FDTD_G double* trace = a.trace;
  const FDTD_G T* dummy = a.mie;  // a valid address for loads whose value is not used
  for (long u = u0; u < u1;) {
    const int t = (int)(u / nzi);
    const long uend = min(u1, (long)(t + 1) * nzi);
    const int ka = 1 + (int)(u - (long)t * nzi), kb = 1 + (int)(uend - (long)t * nzi);
    u = uend;
    const int i0 = 1 + (t % ntx) * (kSCols - 1), j0 = 1 + (t / ntx) * (kSRows - 1);
    const int xi = i0 + cx, yj = j0 + ry;
    const int xc = min(xi, nx - 1), yc = min(yj, ny - 1);
    const bool ecell = xi <= nx - 2 && yj <= ny - 2;
    const bool own = ecell && cx < kSCols - 1 && ry < kSRows - 1;
    const unsigned co = (unsigned)((long)xc + sy * (long)yc) * (unsigned)sizeof(T);
    const int sxi = slab_index(xc, a.slab_a[0], a.slab_b[0]), syi = slab_index(yc, a.slab_a[1], a.slab_b[1]);
    // per-thread x and y coefficients of both field types
    const T fx[2] = {a.cf[0][0][xc], a.cf[1][0][xc]}, fy[2] = {a.cf[0][1][yc], a.cf[1][1][yc]};
    T gx[2], bx[2], cxc[2], gy[2], by[2], cy[2];
#pragma unroll
    for (int ft = 0; ft < 2; ft++) {
      gx[ft] = sxi >= 0 ? a.cg[ft][0][xc] : T(0);
      bx[ft] = sxi >= 0 ? a.cb[ft][0][xc] : T(0);
      cxc[ft] = sxi >= 0 ? a.cc[ft][0][xc] : T(0);
      gy[ft] = syi >= 0 ? a.cg[ft][1][yc] : T(0);
      by[ft] = syi >= 0 ? a.cb[ft][1][yc] : T(0);
      cy[ft] = syi >= 0 ? a.cc[ft][1][yc] : T(0);
    }
    // a psi value (0 .. 2: slab axis) that a cell may need: loaded whatever
    // the cell needs, from the cell's slab entry or, where it has none, from a
    // valid address, so that the gather has no branches (a load inside a
    // branch makes the compiler wait for every memory operation at the join)
    auto psi_ld = [&](const FDTD_G T* base, bool on, long o) -> T {
      const FDTD_G T* p = base ? (on ? base + o : base) : dummy;
      return *p;
    };
    // gather the cell's inputs: old E and E-field psi of plane kE, H-field psi of plane kH
    auto gather = [&](int kE, int kH, SPre& p) {
      const unsigned so = (unsigned)kE * plane_bytes;
#pragma unroll
      for (int q = 0; q < 3; q++) p.eo[q] = buf_ld<T>(rE[q], co, so);
#pragma unroll
      for (int ft = 0; ft < 2; ft++) {
        const int kk = ft == 0 ? kE : max(kH, 1);
        float(&c)[3][2] = ft == 0 ? p.pe : p.ph;
        const bool use = ft == 0 ? ecell : own;
        const int szi = slab_index(kk, a.slab_a[2], a.slab_b[2]);
        const long ox = slab_off(a, 0, max(sxi, 0), xc, yc, kk), oy = slab_off(a, 1, max(syi, 0), xc, yc, kk),
                   oz = slab_off(a, 2, max(szi, 0), xc, yc, kk);
        c[0][0] = psi_ld(a.psi[ft][1][0], use && sxi >= 0, ox);
        c[0][1] = psi_ld(a.psi[ft][2][0], use && sxi >= 0, ox);
        c[1][0] = psi_ld(a.psi[ft][0][1], use && syi >= 0, oy);
        c[1][1] = psi_ld(a.psi[ft][2][1], use && syi >= 0, oy);
        c[2][0] = psi_ld(a.psi[ft][0][2], use && szi >= 0, oz);
        c[2][1] = psi_ld(a.psi[ft][1][2], use && szi >= 0, oz);
      }
    };
    // the update inputs of field type ft at plane kk: the gathered psi, the
    // thread's x and y coefficients, the plane's z coefficients, the sources
    auto pre_of = [&](int ft, int kk, const float (&psi)[3][2]) {
      CellPre<T> c;
#pragma unroll
      for (int J = 0; J < 3; J++) {
        c.psi[J][0] = psi[J][0];
        c.psi[J][1] = psi[J][1];
      }
      c.g[0] = gx[ft];
      c.b[0] = bx[ft];
      c.c[0] = cxc[ft];
      c.g[1] = gy[ft];
      c.b[1] = by[ft];
      c.c[1] = cy[ft];
      c.g[2] = L.zc[ft][1][kk];
      c.b[2] = L.zc[ft][2][kk];
      c.c[2] = L.zc[ft][3][kk];
      c.src = L.src;
      c.nsrc = L.nsrc;
      return c;
    };
    T en[3];  // new E of the cell's current E plane
    // new E of plane kk from the gathered inputs p and the plane's LDS data
    auto e_plane = [&](int kk, bool store, const SPre& p) {
      en[0] = p.eo[0];
      en[1] = p.eo[1];
      en[2] = p.eo[2];
      if (ecell && kk <= nz - 2) {
        const float* h1 = L.h[kk % kSSlots];
        const float* h0 = L.h[(kk - 1) % kSSlots];
        const int mid = L.id[kk % kSSlots][ry * kSCols + cx];
        HStencil<T> h;
        h.hx = h1[(ry + 1) * kSCols + cx];
        h.hx_ym = h1[ry * kSCols + cx];
        h.hy = h1[Gm::oHy + ry * (kSCols + 1) + cx + 1];
        h.hy_xm = h1[Gm::oHy + ry * (kSCols + 1) + cx];
        h.hz = h1[Gm::oHz + (ry + 1) * (kSCols + 1) + cx + 1];
        h.hz_ym = h1[Gm::oHz + ry * (kSCols + 1) + cx + 1];
        h.hz_xm = h1[Gm::oHz + (ry + 1) * (kSCols + 1) + cx];
        h.hx_zm = h0[(ry + 1) * kSCols + cx];
        h.hy_zm = h0[Gm::oHy + ry * (kSCols + 1) + cx + 1];
        const CellPre<T> pre = pre_of(0, kk, p.pe);
        const long c = (long)xi + sy * (long)yj + sz * (long)kk;
        const bool w_own = own && store;
        T o[3];
        const T ie = L.m.ie[mid];  // grids with a design box run the two-pass schedule
        if (tfsf_lds)
          e_cell_3d<T, DISP, false, 3, const double*>(a, step, &L.inc[0][0], xi, yj, kk, c, h, p.eo[0], p.eo[1],
                                                          p.eo[2], mid, ie, fx[0], fy[0], L.zc[0][0][kk], w_own, o,
                                                          &pre);
        else
          e_cell_3d<T, DISP, false, 3>(a, step, a.incH + (long)s * a.auxL, xi, yj, kk, c, h, p.eo[0], p.eo[1],
                                          p.eo[2], mid, ie, fx[0], fy[0], L.zc[0][0][kk], w_own, o, &pre);
        en[0] = o[0];
        en[1] = o[1];
        en[2] = o[2];
        if (w_own) {
          a.Eo[0][c] = o[0];
          a.Eo[1][c] = o[1];
          a.Eo[2][c] = o[2];
        }
      }
    };
    auto publish = [&]() {
      const int x = ry * kSCols + cx;
      L.en[0][x] = en[0];
      L.en[1][x] = en[1];
      L.en[2][x] = en[2];
    };
    T ek[3];
    // plane k of the march: new E of plane k+1 from cur (gathered for planes
    // k+1 / k), new H of plane k; the inputs of planes k+3 / k+2 are gathered
    // into nxt, two planes ahead of their use (three sets rotate without copies)
    auto plane = [&](int k, const SPre& cur, SPre& nxt) {
      s_bar_lds();  // B1: old H and ids of plane k+1 landed, new E of plane k published
      // (near the range's end it gathers planes already read: nothing uses
      // them, and an unconditional gather keeps the rotation a renaming)
      gather(min(k + 3, kb), min(k + 2, kb - 1), nxt);
      e_plane(k + 1, k + 1 < kb, cur);  // the halo plane when k+1 == kb: computed, not stored
      if (own) {  // new H of plane k
        const float* hk = L.h[k % kSSlots];
        const int midk = L.id[k % kSSlots][ry * kSCols + cx];
        const long c = (long)xi + sy * (long)yj + sz * (long)k;
        EStencil<T> e;
        e.ex = ek[0];
        e.ey = ek[1];
        e.ez = ek[2];
        e.ex_yp = L.en[0][(ry + 1) * kSCols + cx];
        e.ez_yp = L.en[2][(ry + 1) * kSCols + cx];
        e.ey_xp = L.en[1][ry * kSCols + cx + 1];
        e.ez_xp = L.en[2][ry * kSCols + cx + 1];
        e.ex_zp = en[0];
        e.ey_zp = en[1];
        const T hx = hk[(ry + 1) * kSCols + cx];
        const T hy = hk[Gm::oHy + ry * (kSCols + 1) + cx + 1];
        const T hz = hk[Gm::oHz + (ry + 1) * (kSCols + 1) + cx + 1];
        const CellPre<T> pre = pre_of(1, k, cur.ph);
        T o[3];
        if (tfsf_lds)
          h_cell_3d<T, false, 3, const double*>(a, step, &L.inc[1][0], xi, yj, k, c, e, hx, hy, hz,
                                                   L.m.im[midk], fx[1], fy[1], L.zc[1][0][k], o, &pre);
        else
          h_cell_3d<T, false, 3>(a, step, a.incE + (long)s * a.auxL, xi, yj, k, c, e, hx, hy, hz, L.m.im[midk],
                                    fx[1], fy[1], L.zc[1][0][k], o, &pre);
        a.Ho[0][c] = o[0];
        a.Ho[1][c] = o[1];
        a.Ho[2][c] = o[2];
        for (int q = 0; q < nprobe; q++)  // the probes of the cell (dft_probe's values, table in LDS)
          if (L.probe[q].cell == c) {
            const int comp = L.probe[q].comp;
            const T v = L.probe[q].field == 1 ? pick3(e.ex, e.ey, e.ez, comp) : pick3(o[0], o[1], o[2], comp);
            trace[step * nprobe + q] = (double)v;
          }
      }
      s_bar_lds();  // B2: every read of the published plane done
      ek[0] = en[0];
      ek[1] = en[1];
      ek[2] = en[2];
      publish();
    };
    SPre P0, A, B, C;
    gather(ka, 0, P0);                           // the prologue's E plane
    gather(ka + 1, ka, A);                       // plane ka of the march
    if (ka + 1 < kb) gather(ka + 2, ka + 1, B);  // plane ka+1
    s_bar_lds();  // B0 (the producers wait for planes ka-1 and ka)
    e_plane(ka, true, P0);
    ek[0] = en[0];
    ek[1] = en[1];
    ek[2] = en[2];
    publish();
    // whole groups of three in the loop (skipped calls inside it would make
    // the loop carry gathered values from several paths: register copies
    // that wait for loads in flight), the rest after it
    int k = ka;
    for (; k + 2 < kb; k += 3) {
      plane(k, A, C);
      plane(k + 1, B, A);
      plane(k + 2, C, B);
    }
    if (k < kb) plane(k, A, C);
    if (k + 1 < kb) plane(k + 1, B, A);
    s_bar_lds();  // the range's end
  }
}

// ---------------------------------------------------------------- z-marching fused pass
// One pass per step.  A workgroup of 64 x TY threads holds a tile of columns
// (i0 - 1 + l, j0 + r), l < 64, r < TY, and marches z over a chunk of planes,
// one cell per thread.  At plane k it computes the new E of its cell from the
// old H (read from memory, as the two passes read it), passes that E to its
// neighbours through LDS, and after one barrier computes the new H of plane
// k-1 from the new E of its own cell at k-1 and k and of its +x, +y
// neighbours at k-1.  The tile's last column and last row compute E only:
// the halo the owned cells' H needs (so the tile owns 63 x (TY - 1)
// columns); the chunk's last plane is computed for E only, as the next
// chunk's first plane.  Old state is only read and new state only written
// (the second copy), so the halo work never races, and the cell arithmetic
// is e_cell_3d / h_cell_3d: the fields are bitwise those of the two passes.
// Three LDS buffers of new E: a thread one plane ahead of another (they are
// never further apart: one barrier per plane) writes the buffer the other
// no longer reads.
template <int N>
using IC = std::integral_constant<int, N>;

constexpr int kZCols = 64;
template <typename T, int TY, int DISP>
__global__ __launch_bounds__(kZCols * TY) void fusedz_3d(const StepArgs<T>* __restrict__ ap, int kc, int s, int ntx,
                                                        int nty, int nkc) {
  const StepArgs<T>& a = *ap;
  __shared__ MatLds<T> sm;
  __shared__ T eb[3][3][TY][kZCols];  // [buffer][component][row][lane]
  stage_materials(a, sm);
  const int t = xcd_tile(blockIdx.x, ntx * nty * nkc);
  const int bx = t % ntx, by = (t / ntx) % nty, bz = t / (ntx * nty);
  const int l = threadIdx.x, r = threadIdx.y;
  const int i = 1 + bx * (kZCols - 1) + l, j = 1 + by * (TY - 1) + r;
  const int k0 = 1 + bz * kc, k1 = min(k0 + kc, a.nz - 1);  // H planes [k0, k1), E planes [k0, k1]
  const bool incol = i <= a.nx - 2 && j <= a.ny - 2;       // an interior column (E is updated there)
  const bool owner = incol && l < kZCols - 1 && r < TY - 1;
  const bool inside = i <= a.nx - 1 && j <= a.ny - 1;     // a grid column: its old E and H exist
  const long sy = a.sy, sz = a.sz;
  const long step = a.step0 + s;
  const FDTD_G double* incH = a.incH + (long)s * a.auxL;
  const FDTD_G double* incE = a.incE + (long)s * a.auxL;
  const T fxE = a.cf[0][0][incol ? i : 1], fyE = a.cf[0][1][incol ? j : 1];
  const T fxH = a.cf[1][0][incol ? i : 1], fyH = a.cf[1][1][incol ? j : 1];
  const long c0 = (long)i + sy * (long)j;
  T ep[3] = {T(0), T(0), T(0)};  // new E of the cell at the previous plane
  // the old state of a plane: the cell's H and E, its id and the -x, -y H
  // neighbours; loaded one plane ahead, so a plane's loads are in flight
  // while the previous plane computes
  struct Plane {
    T h[3], e[3], hz_ym, hx_ym, hz_xm, hy_xm;
    int mid;
  };
  auto load = [&](int k, Plane& q) {
    const long c = c0 + sz * (long)k;
    if (inside) {
      q.h[0] = a.H[0][c];
      q.h[1] = a.H[1][c];
      q.h[2] = a.H[2][c];
      q.e[0] = a.E[0][c];
      q.e[1] = a.E[1][c];
      q.e[2] = a.E[2][c];
      q.mid = a.mat[c];
    }
    if (incol) {
      q.hz_ym = a.H[2][c - sy];
      q.hx_ym = a.H[0][c - sy];
      q.hz_xm = a.H[2][c - 1];
      q.hy_xm = a.H[1][c - 1];
    }
  };
  // three plane slots, plane k in slot (k - k0 + 1) mod 3, over a three times
  // unrolled loop: plane k + 1 loads into the slot of plane k - 2 while plane
  // k - 1's old H and id are read in place, so the loop carries no copy of a
  // plane whose loads are still in flight (a copy would wait for them and
  // leave nothing to overlap with the next plane's arithmetic)
  Plane P[3];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    P[q].h[0] = P[q].h[1] = P[q].h[2] = P[q].e[0] = P[q].e[1] = P[q].e[2] = T(0);
    P[q].hz_ym = P[q].hx_ym = P[q].hz_xm = P[q].hy_xm = T(0);
    P[q].mid = 0;
  }
  // old H of the cell at the plane below: the z-1 neighbour of plane k0's E
  if (inside) {
    P[0].h[0] = a.H[0][c0 + sz * (long)(k0 - 1)];
    P[0].h[1] = a.H[1][c0 + sz * (long)(k0 - 1)];
  }
  load(k0, P[1]);
  auto it = [&](int k, auto sk) {
    constexpr int A = decltype(sk)::value, M1 = (A + 2) % 3, N1 = (A + 1) % 3;
    Plane& cur = P[A];
    const Plane& prv = P[M1];  // plane k - 1
    const long c = c0 + sz * (long)k;
    load(k < k1 ? k + 1 : k1, P[N1]);  // (the last plane loads one nothing reads)
    T en[3] = {cur.e[0], cur.e[1], cur.e[2]};
    if (incol && k <= a.nz - 2) {
      HStencil<T> hs;
      hs.hx = cur.h[0];
      hs.hy = cur.h[1];
      hs.hz = cur.h[2];
      hs.hz_ym = cur.hz_ym;
      hs.hx_ym = cur.hx_ym;
      hs.hz_xm = cur.hz_xm;
      hs.hy_xm = cur.hy_xm;
      hs.hy_zm = prv.h[1];
      hs.hx_zm = prv.h[0];
      const bool own = owner && k < k1;
      T o[3];
      e_cell_3d<T, DISP>(a, step, incH, i, j, k, c, hs, cur.e[0], cur.e[1], cur.e[2], cur.mid,
                         cell_inv_eps(a, sm.ie, cur.mid, i, j, k), fxE, fyE, a.cf[0][2][k], own, o);
      en[0] = o[0];
      en[1] = o[1];
      en[2] = o[2];
      if (own) {
        a.Eo[0][c] = o[0];
        a.Eo[1][c] = o[1];
        a.Eo[2][c] = o[2];
      }
    }
    const int b = k % 3;
    eb[b][0][r][l] = en[0];
    eb[b][1][r][l] = en[1];
    eb[b][2][r][l] = en[2];
    s_bar_lds();  // LDS only: the loads and stores in flight need not land first
    if (owner && k > k0) {
      // new H of plane k-1
      const int bp = (k - 1) % 3;
      const long cp = c - sz;
      EStencil<T> es;
      es.ex = ep[0];
      es.ey = ep[1];
      es.ez = ep[2];
      es.ez_xp = eb[bp][2][r][l + 1];
      es.ey_xp = eb[bp][1][r][l + 1];
      es.ez_yp = eb[bp][2][r + 1][l];
      es.ex_yp = eb[bp][0][r + 1][l];
      es.ey_zp = en[1];
      es.ex_zp = en[0];
      T o[3];
      h_cell_3d(a, step, incE, i, j, k - 1, cp, es, prv.h[0], prv.h[1], prv.h[2], sm.im[prv.mid], fxH, fyH,
                a.cf[1][2][k - 1], o);
      a.Ho[0][cp] = o[0];
      a.Ho[1][cp] = o[1];
      a.Ho[2][cp] = o[2];
      if (a.nprobe) dft_probe(a, step, cp, i, j, k - 1, ep, o);
    }
    ep[0] = en[0];
    ep[1] = en[1];
    ep[2] = en[2];
  };
  int k = k0;
  for (; k + 2 <= k1; k += 3) {
    it(k, IC<1>{});
    it(k + 1, IC<2>{});
    it(k + 2, IC<0>{});
  }
  if (k <= k1) it(k, IC<1>{});
  if (k + 1 <= k1) it(k + 1, IC<2>{});
}

// ---------------------------------------------------------------- time-blocked 3-D pass
// Two steps per launch (nst = 2; nst = 1 runs the first step only).  A
// workgroup of 64 x TY threads holds the region columns (x0 - 2 + l,
// y0 - 2 + r), l < 64, r < TY, one cell per thread per plane, and marches z
// over a chunk of planes [k0, k1) with the four half steps skewed by one
// plane each: at iteration p it computes
//   E2 at plane p - 2 (from H1 of planes p - 2 and p - 3),
//   H2 at plane p - 3 (from E2 of planes p - 3 and p - 2),
//   E1 at plane p     (from the old H of planes p and p - 1),
//   H1 at plane p - 1 (from E1 of planes p - 1 and p).
// E reads its -x, -y neighbours and H its +x, +y ones, so each half step
// leaves one more column / row on one side without exact inputs: E1 is exact
// on l >= 1, H1 on [1, 62], E2 on [2, 62], H2 on the core [2, 61] (rows
// likewise), and the core's 60 x (TY - 4) columns hold both steps' exact
// values.  The core writes them, with its CPML psi, to the second copy of the
// state.  Every lane of a row computes each half step its row takes part in
// (rows outside a half step's range skip it, a wave-uniform branch); the
// values of the lanes outside the exact range are never read by an exact
// one.  Cells outside the interior keep their old values (a select).  x
// neighbours come from the adjacent lane (a wave is one region row, DPP wave
// shifts), y neighbours through LDS, written in the iteration before the one
// that reads them: one barrier per plane.  The first step's CPML psi of a
// cell stays in registers until the second step's update of the cell; the
// planes' values rotate through four register slots (the loop is unrolled
// four times, so the rotation is register renaming).  Fields are addressed
// by 32-bit byte offsets from their bases (one offset register per thread
// and plane).  The cell arithmetic is e_cell_3d / h_cell_3d's, operation for
// operation, so the fields, psi, DFT samples and probes are bitwise those of
// the two passes.  Grids with dispersive cells, boundary conditions, E / H
// sources, single planes, complex fields, a design box, recorded design
// history or arrays of 2 GiB or more take the two passes.
// Everything the march reads besides the fields, ids and psi is staged in LDS
// first (z coefficients of the chunk, the two steps' incident rows and source
// values, the DFT monitors of each plane): a uniform value read from global
// memory inside the loop would be a vector load, and waiting for it would
// wait for the prefetched planes too.
constexpr int kTbCols = 64;
constexpr int kTbCore = kTbCols - 4;
constexpr int kTbMaxKc = 64;     // planes per workgroup at most (FP32; FP64: kTbMaxKc64)
constexpr int kTbMaxKc64 = 32;
constexpr int kTbRowMax = 1024;  // incident-line cells at most
constexpr int kDppWaveShr1 = 0x138;  // lane l receives lane l - 1
constexpr int kDppWaveShl1 = 0x130;  // lane l receives lane l + 1

template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// an element at a 32-bit byte offset from an array's base (base in scalar
// registers, offset in a vector register: one address register for all arrays)
template <typename T>
__device__ __forceinline__ T ldo(const FDTD_G T* base, unsigned off) {
  return *(const FDTD_G T*)((const FDTD_G char*)base + off);
}
template <typename T>
__device__ __forceinline__ void sto(FDTD_G T* base, unsigned off, T v) {
  *(FDTD_G T*)((FDTD_G char*)base + off) = v;
}

// DFT samples of one cell after a step (the time-blocked form of
// dft_sample_block: the same slots), fld 0: E values, 1: H values
template <typename T>
__device__ FDTD_COLD void tb_sample(const StepArgs<T>& a, long step, int i, int j, int k, int fld, const T v[3]) {
  for (int q = 0; q < a.ndft; q++) {
    const DftDev<T>& d = a.dft[q];
    if (i < d.r[0] || i >= d.r[1] || j < d.r[2] || j >= d.r[3] || k < d.r[4] || k >= d.r[5]) continue;
    const int wx = d.r[1] - d.r[0], wy = d.r[3] - d.r[2];
    const long lq = (long)(i - d.r[0]) + (long)wx * ((long)(j - d.r[2]) + (long)wy * (k - d.r[4]));
    const long slot0 = (long)((int)(step % kDftBlock) * d.ncomp);
    int slot = 0;
    for (int comp = 0; comp < 6; comp++) {
      if (!((d.mask >> comp) & 1)) continue;
      if ((comp < 3) == (fld == 0)) d.smp[(slot0 + slot) * d.ncells + lq] = v[comp % 3];
      slot++;
    }
  }
}


template <typename T, int TY>
__global__ __launch_bounds__(kTbCols * TY) void tb_3d(const StepArgs<T>* __restrict__ ap, int kc, int s0, int nst,
                                                     int ntx, int nty, int nkc) {
#pragma clang fp contract(off)
  const StepArgs<T>& a = *ap;
  __shared__ MatLds<T> sm;
  // y-neighbour exchange [buffer][quantity][row][lane]: H0(p+1) hz, hx; E1(p)
  // ez, ex; H1(p-1) hz, hx; E2(p-2) ez, ex (written at iteration p)
  __shared__ T xs[2][8][TY][kTbCols];
  __shared__ T zc[2][4][kTbMaxKc + 5];   // z coefficients dt F/dx, dt G, B, C/dx of planes k0 - 2 .. k1 + 2
  __shared__ double rows[4][kTbRowMax];  // incident rows: aux H of steps 1, 2; aux E of steps 1, 2
  __shared__ SrcVal<T> ssrc[2][kMaxSrc];  // the two steps' source values (sources still running)
  __shared__ int nss[2];
  __shared__ unsigned char dpl[kTbMaxKc + 5];  // DFT monitors holding each plane (bit per monitor)
  // TFSF x-face terms of every plane of the chunk, computed by the whole
  // workgroup before the march (a few terms per thread) instead of by every
  // wave holding a face lane at every half step: [plane][field type][step]
  // [face side][row][I - 1], and whether the component's face cell lies
  // inside the box's face (the term is added)
  constexpr int kTab = (sizeof(T) == 4 ? kTbMaxKc : kTbMaxKc64) + 5;
  __shared__ T xt[kTab][2][2][2][TY][2];
  __shared__ unsigned char xin[kTab][2][2][2][TY][2];
  __shared__ T xfc[2][2];  // dt F/dx along x at the two x faces [field type][side]
  const int t = xcd_tile(blockIdx.x, ntx * nty * nkc);
  const int bx = t % ntx, by = (t / ntx) % nty, bz = t / (ntx * nty);
  // a wave is one row of the region: the row index is wave-uniform
  const int l = threadIdx.x, r = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const int i = bx * kTbCore - 1 + l, j = by * (TY - 4) - 1 + r;
  const int k0 = 1 + bz * kc, k1 = min(k0 + kc, a.nz - 1);
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int tid = l + kTbCols * r;
  // ---- stage the launch's tables
  const int kz0 = k0 - 2, nzs = min(k1 + 2, nz - 1) - kz0 + 1;
  for (int q = tid; q < 8 * nzs; q += kTbCols * TY) {
    const int ft = q / (4 * nzs), w = (q / nzs) % 4, z = q % nzs, k = kz0 + z;
    const FDTD_G T* src = w == 0 ? a.cf[ft][2] : (w == 1 ? a.cg[ft][2] : (w == 2 ? a.cb[ft][2] : a.cc[ft][2]));
    zc[ft][w][z] = k >= 0 ? src[k] : T(0);
  }
  for (int q = tid; q < nzs; q += kTbCols * TY) {
    unsigned m = 0;
    for (int d = 0; d < a.ndft; d++)
      if (kz0 + q >= a.dft[d].r[4] && kz0 + q < a.dft[d].r[5]) m |= 1u << d;
    dpl[q] = (unsigned char)m;
  }
  if (a.tfsf)
    for (int q = tid; q < a.auxL * nst; q += kTbCols * TY) {
      const int u = q / a.auxL, x = q - u * a.auxL;
      rows[u][x] = a.incH[(long)(s0 + u) * a.auxL + x];
      rows[2 + u][x] = a.incE[(long)(s0 + u) * a.auxL + x];
    }
  if (a.tfsf && tid < 4) xfc[tid >> 1][tid & 1] = a.cf[tid >> 1][0][min(max(a.box[tid & 1], 0), a.nx - 1)];
  if (tid < nst) {
    const long step = a.step0 + s0 + tid;
    int n = 0;
    for (int q = 0; q < a.nsrc; q++)  // add_sources adds a source while step < its table length
      if (step < a.src[q].n)
        ssrc[tid][n++] = SrcVal<T>{a.src[q].cell, a.src[q].field, a.src[q].comp, (T)a.src[q].table[step]};
    nss[tid] = n;
  }
  stage_materials(a, sm);  // (its barrier also covers the tables above)

  // ---- the thread's column
  const bool incol = i >= 1 && j >= 1 && i <= nx - 2 && j <= ny - 2;  // an interior column (updated)
  const bool own = incol && l >= 2 && l <= kTbCols - 3 && r >= 2 && r <= TY - 3;
  const bool two = nst > 1;
  const int rm = r > 0 ? r - 1 : 0, rp = r < TY - 1 ? r + 1 : TY - 1;
  // columns off the grid load (and compute from) the nearest grid column:
  // nothing exact reads them
  const int ic = min(max(i, 0), nx - 1), jc = min(max(j, 0), ny - 1);
  const unsigned c0 = (unsigned)ic + (unsigned)nx * (unsigned)jc;  // cell index of the column's plane 0
  const unsigned pstride = (unsigned)nx * (unsigned)ny;
  const unsigned vo = c0 * (unsigned)sizeof(T), pb = pstride * (unsigned)sizeof(T);
  const long st1 = a.step0 + s0, st2 = st1 + 1;
  const int ci = incol ? i : 1, cj = incol ? j : 1;
  const T fxE = a.cf[0][0][ci], fyE = a.cf[0][1][cj], fxH = a.cf[1][0][ci], fyH = a.cf[1][1][cj];
  const int sxi = incol ? slab_index(i, a.slab_a[0], a.slab_b[0]) : -1;
  const int syi = incol ? slab_index(j, a.slab_a[1], a.slab_b[1]) : -1;
  T gxE = T(0), bxE = T(0), cxE = T(0), gyE = T(0), byE = T(0), cyE = T(0);
  T gxH = T(0), bxH = T(0), cxH = T(0), gyH = T(0), byH = T(0), cyH = T(0);
  if (sxi >= 0) {
    gxE = a.cg[0][0][i];
    bxE = a.cb[0][0][i];
    cxE = a.cc[0][0][i];
    gxH = a.cg[1][0][i];
    bxH = a.cb[1][0][i];
    cxH = a.cc[1][0][i];
  }
  if (syi >= 0) {
    gyE = a.cg[0][1][j];
    byE = a.cb[0][1][j];
    cyE = a.cc[0][1][j];
    gyH = a.cg[1][1][j];
    byH = a.cb[1][1][j];
    cyH = a.cc[1][1][j];
  }
  // psi byte offsets: x slab (sxi + S0 (j + ny k)), y slab (i + nx (syi + S1 k)),
  // z slab (i + nx (j + ny szi)); the k parts per plane
  const unsigned oxb = sxi >= 0 ? ((unsigned)sxi + (unsigned)a.slab_S[0] * (unsigned)j) * (unsigned)sizeof(T) : 0u;
  const unsigned oyb = syi >= 0 ? ((unsigned)i + (unsigned)nx * (unsigned)syi) * (unsigned)sizeof(T) : 0u;
  const unsigned kxb = (unsigned)a.slab_S[0] * (unsigned)ny * (unsigned)sizeof(T);
  const unsigned kyb = (unsigned)nx * (unsigned)a.slab_S[1] * (unsigned)sizeof(T);
  // a column holding a source cell adds the staged source values; one holding
  // a probe cell records it; DFT monitors whose x, y range holds the column
  bool scol = false, pcol = false;
  for (int q = 0; q < a.nsrc; q++) scol = scol || (incol && (unsigned)(a.src[q].cell % (long)pstride) == c0);
  for (int q = 0; q < a.nprobe; q++) pcol = pcol || (own && (unsigned)(a.probe[q].cell % (long)pstride) == c0);
  unsigned dcol = 0;
  for (int q = 0; q < a.ndft; q++)
    if (own && i >= a.dft[q].r[0] && i < a.dft[q].r[1] && j >= a.dft[q].r[2] && j < a.dft[q].r[3]) dcol |= 1u << q;
  // TFSF: the x faces' lanes add the workgroup's precomputed terms; y-face
  // rows and z-face planes compute theirs in the cell update (every lane of
  // such a wave is a face cell)
  const int lf0 = a.tfsf ? a.box[0] - (bx * kTbCore - 1) : -1, lf1 = a.tfsf ? a.box[1] - (bx * kTbCore - 1) : -1;
  const bool xf0 = lf0 >= 0 && lf0 < kTbCols, xf1 = lf1 >= 0 && lf1 < kTbCols && a.box[1] != a.box[0];
  const bool xfaces = xf0 || xf1;
  const int xside = (xf0 && l == lf0) ? 0 : ((xf1 && l == lf1) ? 1 : -1);
  const bool tfy = a.tfsf && incol && (j == a.box[2] || j == a.box[3]);
  // wave-uniform guards of the lane-dependent branches (most waves take none:
  // the uniform test costs a scalar branch, the lane mask a save / restore)
  const bool wx = __ballot(sxi >= 0) != 0, wy = __ballot(syi >= 0) != 0, wsrc = __ballot(scol) != 0;
  const bool wxf = __ballot(xside >= 0) != 0, wty = __ballot(tfy) != 0;
  const bool wd = __ballot(dcol != 0) != 0, wp = __ballot(pcol) != 0;
  // the half steps this row takes part in
  const bool doE1 = r >= 1, doH1 = r >= 1 && r <= TY - 2, doE2 = two && r >= 2 && r <= TY - 2,
             doH2 = two && r >= 2 && r <= TY - 3;
  // the planes each half step covers
  const int e1hi = min(k1 + 1, nz - 2), h1lo = max(1, k0 - 1), h1hi = min(k1, nz - 2), e2hi = min(k1, nz - 2);
  const int sza = a.slab_a[2], szb = a.slab_b[2];
  const bool tz = a.tfsf != 0;
  const int bz4 = a.box[4], bz5 = a.box[5];
  // the x-face terms: tfsf_cell's J = 0 terms, operation for operation
  auto xterm = [&](int x) {
    const int c = x & 1, rr = (x >> 1) % TY, sd = (x / (2 * TY)) & 1, u = (x / (4 * TY)) & 1,
              ft = (x / (8 * TY)) & 1, z = x / (16 * TY);
    if (!(sd == 0 ? xf0 : xf1)) return;
    const int ci3[3] = {a.box[sd], by * (TY - 4) - 1 + rr, kz0 + z};
    const int I = 1 + c, K = 2 - c;
    bool inside = true;
#pragma unroll
    for (int Tn = 1; Tn < 3; Tn++) {
      const bool half = ft == 0 ? (Tn != I) : (Tn == I);
      const int lo = half ? a.box[2 * Tn] + 1 : a.box[2 * Tn];
      const int hi = half ? a.box[2 * Tn + 1] : a.box[2 * Tn + 1] - 1;
      if (ci3[Tn] < lo || ci3[Tn] > hi) inside = false;
    }
    double proj = 0.0;
#pragma unroll
    for (int dd = 0; dd < 3; dd++) {
      const double off = ft == 0 ? (dd == K ? -0.5 : 0.0) : (dd != K ? -0.5 : 0.0);
      const double loc = fmadd(off, a.dx, ((double)ci3[dd] - a.origin[dd]) * a.dx);
      proj = fmadd(a.kvec[dd], loc, proj);
    }
    const double inc = interp_row((const double*)rows[(ft ? 2 : 0) + u], a.auxL, proj, a.dx, ft == 0);
    const double pv = ft == 0 ? a.polH[K] : a.pol[K];
    const double f = (double)xfc[ft][sd];
    const double term = lc3(I, 0) * f * (sd == 0 ? -1.0 : 1.0) * inc * pv;
    xt[z][ft][u][sd][rr][c] = (T)(ft == 0 ? term : -term);
    xin[z][ft][u][sd][rr][c] = inside ? 1 : 0;
  };
  if (xfaces) {
    for (int x = tid; x < nzs * 16 * TY; x += kTbCols * TY) xterm(x);
    __syncthreads();
  }

  // ---- the cell updates (e_cell_3d / h_cell_3d's arithmetic)
  auto add_src = [&](int u, int field, int q, T d[3]) {
    const long c = (long)c0 + (long)pstride * q;
    for (int s = 0; s < nss[u]; s++) {
      const SrcVal<T>& v = ssrc[u][s];
      if (v.cell == c && v.field == field) {
        if (v.comp == 0)
          d[0] += v.v;
        else if (v.comp == 1)
          d[1] += v.v;
        else
          d[2] += v.v;
      }
    }
  };
  // new E of the cell at plane q, step u, from its H stencil; ps: psi in / out
  auto e_upd = [&](int u, int q, T hx, T hy, T hz, T hz_ym, T hz_xm, T hy_xm, T hy_zm, T hx_ym, T hx_zm, T ex,
                   T ey, T ez, T ie, T (&ps)[3][2], T out[3]) {
    const int z = q - kz0;
    const T fz = zc[0][0][z];
    const T dxy = hz - hz_ym;  // Ex, J=y, +1
    const T dxz = hy - hy_zm;  // Ex, J=z, -1
    const T dyz = hx - hx_zm;  // Ey, J=z, +1
    const T dyx = hz - hz_xm;  // Ey, J=x, -1
    const T dzx = hy - hy_xm;  // Ez, J=x, +1
    const T dzy = hx - hx_ym;  // Ez, J=y, -1
    T d[3];
    d[0] = fmadd(fyE, dxy, -(fz * dxz));
    d[1] = fmadd(fz, dyz, -(fxE * dyx));
    d[2] = fmadd(fxE, dzx, -(fyE * dzy));
    if (wsrc && scol) add_src(u, 0, q, d);
    if (wx && sxi >= 0) {
      const T p1 = ps[0][0], p2 = ps[0][1];
      d[1] = fmadd(gxE, p1, d[1]);
      d[2] = fmadd(gxE, p2, d[2]);
      ps[0][0] = fmadd(bxE, p1, -(cxE * dyx));
      ps[0][1] = fmadd(bxE, p2, cxE * dzx);
    }
    if (wy && syi >= 0) {
      const T p0 = ps[1][0], p2 = ps[1][1];
      d[0] = fmadd(gyE, p0, d[0]);
      d[2] = fmadd(gyE, p2, d[2]);
      ps[1][0] = fmadd(byE, p0, cyE * dxy);
      ps[1][1] = fmadd(byE, p2, -(cyE * dzy));
    }
    if (q < sza || q >= szb) {
      const T g = zc[0][1][z], b = zc[0][2][z], cc = zc[0][3][z];
      const T p0 = ps[2][0], p1 = ps[2][1];
      d[0] = fmadd(g, p0, d[0]);
      d[1] = fmadd(g, p1, d[1]);
      ps[2][0] = fmadd(b, p0, -(cc * dxz));
      ps[2][1] = fmadd(b, p1, cc * dyz);
    }
    if (wxf && xside >= 0) {
      if (xin[z][0][u][xside][r][0]) d[1] += xt[z][0][u][xside][r][0];
      if (xin[z][0][u][xside][r][1]) d[2] += xt[z][0][u][xside][r][1];
    }
    if ((tz && (q == bz4 || q == bz5)) || (wty && tfy)) {
      const bool act[3] = {true, true, true};
      const T fn[3] = {fxE, fyE, fz};
      tfsf_cell<T, 3, false, const double*, 1>(a, rows[u], 0, i, j, q, act, act, fn, d);
    }
    if (ie > T(0)) {
      out[0] = fmadd(d[0], ie, ex);
      out[1] = fmadd(d[1], ie, ey);
      out[2] = fmadd(d[2], ie, ez);
    } else {
      out[0] = ex;
      out[1] = ey;
      out[2] = ez;
    }
  };
  auto h_upd = [&](int u, int q, T ex, T ey, T ez, T ez_yp, T ez_xp, T ey_xp, T ey_zp, T ex_yp, T ex_zp, T hx,
                   T hy, T hz, T im, T (&ps)[3][2], T out[3]) {
    const int z = q - kz0;
    const T fz = zc[1][0][z];
    const T dxy = ez_yp - ez;  // Hx, J=y, +1
    const T dxz = ey_zp - ey;  // Hx, J=z, -1
    const T dyz = ex_zp - ex;  // Hy, J=z, +1
    const T dyx = ez_xp - ez;  // Hy, J=x, -1
    const T dzx = ey_xp - ey;  // Hz, J=x, +1
    const T dzy = ex_yp - ex;  // Hz, J=y, -1
    T d[3];
    d[0] = fmadd(fz, dxz, -(fyH * dxy));
    d[1] = fmadd(fxH, dyx, -(fz * dyz));
    d[2] = fmadd(fyH, dzy, -(fxH * dzx));
    if (wsrc && scol) add_src(u, 2, q, d);
    if (wx && sxi >= 0) {
      const T p1 = ps[0][0], p2 = ps[0][1];
      d[1] = fmadd(-gxH, p1, d[1]);
      d[2] = fmadd(-gxH, p2, d[2]);
      ps[0][0] = fmadd(bxH, p1, -(cxH * dyx));
      ps[0][1] = fmadd(bxH, p2, cxH * dzx);
    }
    if (wy && syi >= 0) {
      const T p0 = ps[1][0], p2 = ps[1][1];
      d[0] = fmadd(-gyH, p0, d[0]);
      d[2] = fmadd(-gyH, p2, d[2]);
      ps[1][0] = fmadd(byH, p0, cyH * dxy);
      ps[1][1] = fmadd(byH, p2, -(cyH * dzy));
    }
    if (q < sza || q >= szb) {
      const T g = zc[1][1][z], b = zc[1][2][z], cc = zc[1][3][z];
      const T p0 = ps[2][0], p1 = ps[2][1];
      d[0] = fmadd(-g, p0, d[0]);
      d[1] = fmadd(-g, p1, d[1]);
      ps[2][0] = fmadd(b, p0, -(cc * dxz));
      ps[2][1] = fmadd(b, p1, cc * dyz);
    }
    if (wxf && xside >= 0) {
      if (xin[z][1][u][xside][r][0]) d[1] += xt[z][1][u][xside][r][0];
      if (xin[z][1][u][xside][r][1]) d[2] += xt[z][1][u][xside][r][1];
    }
    if ((tz && (q == bz4 || q == bz5)) || (wty && tfy)) {
      const bool act[3] = {true, true, true};
      const T fn[3] = {fxH, fyH, fz};
      tfsf_cell<T, 3, false, const double*, 1>(a, rows[2 + u], 1, i, j, q, act, act, fn, d);
    }
    out[0] = fmadd(d[0], im, hx);
    out[1] = fmadd(d[1], im, hy);
    out[2] = fmadd(d[2], im, hz);
  };
  // psi of field type ft at plane q (the slabs the cell is in)
  auto load_psi = [&](int ft, int q, T (&ps)[3][2]) {
    if (q < 1 || q > nz - 2) return;
    if (wx && sxi >= 0) {
      const unsigned o = oxb + kxb * (unsigned)q;
      ps[0][0] = ldo(a.psi[ft][1][0], o);
      ps[0][1] = ldo(a.psi[ft][2][0], o);
    }
    if (wy && syi >= 0) {
      const unsigned o = oyb + kyb * (unsigned)q;
      ps[1][0] = ldo(a.psi[ft][0][1], o);
      ps[1][1] = ldo(a.psi[ft][2][1], o);
    }
    const int szi = slab_index(q, sza, szb);
    if (szi >= 0) {
      const unsigned o = vo + pb * (unsigned)szi;
      ps[2][0] = ldo(a.psi[ft][0][2], o);
      ps[2][1] = ldo(a.psi[ft][1][2], o);
    }
  };
  auto store_psi = [&](int ft, int q, const T (&ps)[3][2]) {
    if (wx && sxi >= 0) {
      const unsigned o = oxb + kxb * (unsigned)q;
      sto(a.psio[ft][1][0], o, ps[0][0]);
      sto(a.psio[ft][2][0], o, ps[0][1]);
    }
    if (wy && syi >= 0) {
      const unsigned o = oyb + kyb * (unsigned)q;
      sto(a.psio[ft][0][1], o, ps[1][0]);
      sto(a.psio[ft][2][1], o, ps[1][1]);
    }
    const int szi = slab_index(q, sza, szb);
    if (szi >= 0) {
      const unsigned o = vo + pb * (unsigned)szi;
      sto(a.psio[ft][0][2], o, ps[2][0]);
      sto(a.psio[ft][1][2], o, ps[2][1]);
    }
  };

  // ---- register slots: plane q's values live in slot (q - k0 + 1) mod 4.  Four
  // slots (planes p - 1 .. p + 2 of the old state in flight or in use) make the
  // rotation over a four-times unrolled loop a pure renaming: plane p + 2 loads
  // into the slot of plane p - 2, no longer read, while plane p - 1's old H is
  // still read in place
  struct Ld {
    T e[3], h[3];
    int mid;
  };
  Ld L[4];        // old state of planes p - 1 .. p + 2
  T E1[4][3];     // E1 of planes p, p - 1, p - 2
  T H1[4][3];     // H1 of planes p - 1, p - 2, p - 3
  T E2[4][3];     // E2 of planes p - 2, p - 3
  T PE[4][3][2];  // first-step E psi of planes p - 2, p - 1, p (loaded old psi, then updated)
  T PH[4][3][2];  // first-step H psi of planes p - 3, p - 2, p - 1 (loaded old psi, then updated)
#pragma unroll
  for (int s = 0; s < 4; s++) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      L[s].e[c] = L[s].h[c] = E1[s][c] = H1[s][c] = E2[s][c] = T(0);
      PE[s][c][0] = PE[s][c][1] = PH[s][c][0] = PH[s][c][1] = T(0);
    }
    L[s].mid = 0;
  }
  auto load = [&](int q, Ld& d) {
    const int qc = min(max(q, 0), nz - 1);
    const unsigned o = vo + pb * (unsigned)qc;
    d.h[0] = ldo(a.H[0], o);
    d.h[1] = ldo(a.H[1], o);
    d.h[2] = ldo(a.H[2], o);
    d.e[0] = ldo(a.E[0], o);
    d.e[1] = ldo(a.E[1], o);
    d.e[2] = ldo(a.E[2], o);
    d.mid = ldo(a.mat, c0 + pstride * (unsigned)qc);
  };
  load(k0 - 2, L[3]);
  load(k0 - 1, L[0]);
  load(k0, L[1]);
  load_psi(0, k0 - 1, PE[0]);
  unsigned mids = (unsigned)L[3].mid;  // material ids of planes p - 1, p - 2, p - 3 (bytes 0..2)
  // prologue: the y neighbours of the first iteration's old H
  {
    const int wb = (k0 - 2) & 1;
#pragma unroll
    for (int q = 2; q < 8; q++) xs[wb][q][r][l] = T(0);
    xs[wb][0][r][l] = L[0].h[2];
    xs[wb][1][r][l] = L[0].h[0];
  }
  s_bar_lds();

  // one iteration; S is plane p's slot: plane p - 1 in slot (S + 3) & 3,
  // p - 2 and p + 2 in (S + 2) & 3, p - 3 and p + 1 in (S + 1) & 3
  auto iter = [&](int p, auto sp) {
    constexpr int A = decltype(sp)::value, M1 = (A + 3) & 3, M2 = (A + 2) & 3, M3 = (A + 1) & 3;
    const int rb = (p - 1) & 1, wb = p & 1;
    const int m0 = L[A].mid, m1 = mids & 255, m2 = (mids >> 8) & 255, m3 = (mids >> 16) & 255;
    const T(&h0p)[3] = L[M1].h;  // plane p - 1's old H
    load(p + 2, L[M2]);
    // ---- E2 at plane q2 = p - 2 (second step)
    const int q2 = p - 2;
    if (doE2 && q2 >= k0 && q2 <= e2hi) {
      const T hz_xm = dpp_mov<kDppWaveShr1>(H1[M2][2]), hy_xm = dpp_mov<kDppWaveShr1>(H1[M2][1]);
      T o[3];
      e_upd(1, q2, H1[M2][0], H1[M2][1], H1[M2][2], xs[rb][4][rm][l], hz_xm, hy_xm, H1[M3][1], xs[rb][5][rm][l],
            H1[M3][0], E1[M2][0], E1[M2][1], E1[M2][2], sm.ie[m2], PE[M2], o);
#pragma unroll
      for (int c = 0; c < 3; c++) E2[M2][c] = incol ? o[c] : E1[M2][c];
      if (q2 < k1 && own) {
        const unsigned so = vo + pb * (unsigned)q2;
        sto(a.Eo[0], so, o[0]);
        sto(a.Eo[1], so, o[1]);
        sto(a.Eo[2], so, o[2]);
        store_psi(0, q2, PE[M2]);
        if (wd && dcol && (dcol & dpl[q2 - kz0])) tb_sample(a, st2, i, j, q2, 0, o);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) E2[M2][c] = E1[M2][c];
    }
    // ---- H2 at plane q3 = p - 3 (second step, core rows)
    const int q3 = p - 3;
    if (doH2 && q3 >= k0 && q3 < k1) {
      const T ez_xp = dpp_mov<kDppWaveShl1>(E2[M3][2]), ey_xp = dpp_mov<kDppWaveShl1>(E2[M3][1]);
      T o[3];
      h_upd(1, q3, E2[M3][0], E2[M3][1], E2[M3][2], xs[rb][6][rp][l], ez_xp, ey_xp, E2[M2][1], xs[rb][7][rp][l],
            E2[M2][0], H1[M3][0], H1[M3][1], H1[M3][2], sm.im[m3], PH[M3], o);
      if (own) {
        const unsigned so = vo + pb * (unsigned)q3;
        sto(a.Ho[0], so, o[0]);
        sto(a.Ho[1], so, o[1]);
        sto(a.Ho[2], so, o[2]);
        store_psi(1, q3, PH[M3]);
        if (wd && dcol && (dcol & dpl[q3 - kz0])) tb_sample(a, st2, i, j, q3, 1, o);
        if (wp && pcol) {
          const T ev[3] = {E2[M3][0], E2[M3][1], E2[M3][2]};
          dft_probe(a, st2, (long)c0 + (long)pstride * q3, i, j, q3, ev, o);
        }
      }
    }
    // ---- E1 at plane p (first step)
    if (doE1 && p >= 1 && p <= e1hi) {
      const T hz_xm = dpp_mov<kDppWaveShr1>(L[A].h[2]), hy_xm = dpp_mov<kDppWaveShr1>(L[A].h[1]);
      T o[3];
      e_upd(0, p, L[A].h[0], L[A].h[1], L[A].h[2], xs[rb][0][rm][l], hz_xm, hy_xm, h0p[1], xs[rb][1][rm][l],
            h0p[0], L[A].e[0], L[A].e[1], L[A].e[2], sm.ie[m0], PE[A], o);
#pragma unroll
      for (int c = 0; c < 3; c++) E1[A][c] = incol ? o[c] : L[A].e[c];
      if (p >= k0 && p < k1 && own) {
        if (!two) {
          const unsigned so = vo + pb * (unsigned)p;
          sto(a.Eo[0], so, o[0]);
          sto(a.Eo[1], so, o[1]);
          sto(a.Eo[2], so, o[2]);
          store_psi(0, p, PE[A]);
        }
        if (wd && dcol && (dcol & dpl[p - kz0])) tb_sample(a, st1, i, j, p, 0, o);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) E1[A][c] = L[A].e[c];
    }
    // ---- H1 at plane q1 = p - 1 (first step)
    const int q1 = p - 1;
    if (doH1 && q1 >= h1lo && q1 <= h1hi) {
      const T ez_xp = dpp_mov<kDppWaveShl1>(E1[M1][2]), ey_xp = dpp_mov<kDppWaveShl1>(E1[M1][1]);
      T o[3];
      h_upd(0, q1, E1[M1][0], E1[M1][1], E1[M1][2], xs[rb][2][rp][l], ez_xp, ey_xp, E1[A][1], xs[rb][3][rp][l],
            E1[A][0], h0p[0], h0p[1], h0p[2], sm.im[m1], PH[M1], o);
#pragma unroll
      for (int c = 0; c < 3; c++) H1[M1][c] = incol ? o[c] : h0p[c];
      if (q1 >= k0 && q1 < k1 && own) {
        if (!two) {
          const unsigned so = vo + pb * (unsigned)q1;
          sto(a.Ho[0], so, o[0]);
          sto(a.Ho[1], so, o[1]);
          sto(a.Ho[2], so, o[2]);
          store_psi(1, q1, PH[M1]);
        }
        if (wd && dcol && (dcol & dpl[q1 - kz0])) tb_sample(a, st1, i, j, q1, 1, o);
        if (wp && pcol) {
          const T ev[3] = {E1[M1][0], E1[M1][1], E1[M1][2]};
          dft_probe(a, st1, (long)c0 + (long)pstride * q1, i, j, q1, ev, o);
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) H1[M1][c] = h0p[c];
    }
    // the next iteration's old psi: E at plane p + 1 (the slot of plane p - 3,
    // whose E psi is no longer read), H at plane p (the slot of plane p - 4)
    load_psi(0, p + 1, PE[M3]);
    load_psi(1, p, PH[A]);
    // the y neighbours the next iteration reads
    xs[wb][0][r][l] = L[M3].h[2];
    xs[wb][1][r][l] = L[M3].h[0];
    xs[wb][2][r][l] = E1[A][2];
    xs[wb][3][r][l] = E1[A][0];
    xs[wb][4][r][l] = H1[M1][2];
    xs[wb][5][r][l] = H1[M1][0];
    xs[wb][6][r][l] = E2[M2][2];
    xs[wb][7][r][l] = E2[M2][0];
    s_bar_lds();  // LDS only: the loads and stores in flight need not land first
    mids = ((mids << 8) | (unsigned)m0) & 0xffffffu;
  };
  // whole groups of four iterations: the last group's iterations past k1 + 2
  // update nothing (every half step and store checks its plane range; loads
  // clamp to the grid), and an unconditional body keeps the slot rotation a
  // renaming: skipped iterations would make the loop carry values from several
  // paths, i.e. register copies of loads still in flight, which wait for them
  for (int p = k0 - 1; p <= k1 + 2; p += 4) {
    iter(p, IC<0>{});
    iter(p + 1, IC<1>{});
    iter(p + 2, IC<2>{});
    iter(p + 3, IC<3>{});
  }
}

// ---------------------------------------------------------------- 2-D TM
// Ez, Hx, Hy (include/YeeFields.hpp:123-148); CPML along x and y; in place.
// GEN: the general instance (boundary ghost writes, E / H soft sources,
// single TFSF planes).
template <typename T, bool GEN = false>
__global__ __launch_bounds__(256) void e_pass_tm(const StepArgs<T>* __restrict__ ap, int, int s) {
#pragma clang fp contract(off)
  const StepArgs<T>& a = *ap;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i < 1 || i >= a.nx - 1 || j >= a.ny - 1) return;
  const FDTD_G T* __restrict__ Hx = a.H[0];
  const FDTD_G T* __restrict__ Hy = a.H[1];
  const long step = a.step0 + s;
  const long c = (long)i + a.sy * j;
  const T hx = Hx[c], hy = Hy[c], hy_xm = Hy[c - 1], hx_ym = Hx[c - a.sy];
  T ez = a.E[2][c];
  if (GEN && a.nsrc_eh) {  // the old Ez without the last step's E source value
    T v[3] = {T(0), T(0), ez};
    soft_eh(a, step, 1, c, true, v);
    ez = v[2];
  }
  const int mid = a.mat[c];
  const T ie = cell_inv_eps(a, a.mie, mid, i, j, 0);
  const T dzx = hy - hy_xm;  // J=x, +1
  const T dzy = hx - hx_ym;  // J=y, -1
  T d[3] = {T(0), T(0), T(0)};
  d[2] = fmadd(a.cf[0][0][i], dzx, -(a.cf[0][1][j] * dzy));
  if (a.nsrc) add_sources(a, step, 0, c, d);
  const int sxi = slab_index(i, a.slab_a[0], a.slab_b[0]);
  if (sxi >= 0) {
    const long o = slab_off(a, 0, sxi, i, j, 0);
    const T p = a.psi[0][2][0][o];
    d[2] = fmadd(a.cg[0][0][i], p, d[2]);
    a.psio[0][2][0][o] = fmadd(a.cb[0][0][i], p, a.cc[0][0][i] * dzx);
  }
  const int syi = slab_index(j, a.slab_a[1], a.slab_b[1]);
  if (syi >= 0) {
    const long o = slab_off(a, 1, syi, i, j, 0);
    const T p = a.psi[0][2][1][o];
    d[2] = fmadd(a.cg[0][1][j], p, d[2]);
    a.psio[0][2][1][o] = fmadd(a.cb[0][1][j], p, -(a.cc[0][1][j] * dzy));
  }
  if (a.tfsf && (i == a.box[0] || i == a.box[1] || j == a.box[2] || j == a.box[3])) {
    const bool aF[3] = {false, false, true}, aC[3] = {true, true, false};
    tfsf_cell<T, 2>(a, a.incH + (long)s * a.auxL, 0, i, j, 0, aF, aC, (const T*)nullptr, d);
  }
  if (GEN && a.nplane) plane_cells<T, 2>(a, step, 0, i, j, 0, 4, 3, d);
  T ezn = ez;
  if (ie > T(0)) {
    ezn = fmadd(d[2], ie, ez);
    a.Eo[2][c] = ezn;
  } else if (ie < T(0) && a.Dd[2]) {
    const MatDev<T>& m = a.mats[mid];
    if (m.fluid)
      ezn = fluid_component<T, 2>(a, m, dbox_index(a, i, j, 0), d[2], true);
    else
      ezn = ade_component<T, 2>(a, m, dbox_index(a, i, j, 0), d[2], ez, true);
    a.Eo[2][c] = ezn;
  }
  if (GEN && a.nsrc_eh) {
    T v[3] = {T(0), T(0), ezn};
    soft_eh(a, step, 1, c, false, v);
    ezn = v[2];
    a.Eo[2][c] = ezn;
  }
  if (GEN && a.bc_any) {
    const T v[3] = {T(0), T(0), ezn};
    bc_ghost_write(a, 0, 2, i, j, 0, c, v);
  }
}

template <typename T, bool GEN = false>
__global__ __launch_bounds__(256) void h_pass_tm(const StepArgs<T>* __restrict__ ap, int, int s) {
#pragma clang fp contract(off)
  const StepArgs<T>& a = *ap;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i < 1 || i >= a.nx - 1 || j >= a.ny - 1) return;
  const FDTD_G T* __restrict__ Ez = a.Eo[2];
  const long step = a.step0 + s;
  const long c = (long)i + a.sy * j;
  const T ez = Ez[c], ez_yp = Ez[c + a.sy], ez_xp = Ez[c + 1];
  T hx = a.H[0][c], hy = a.H[1][c];
  if (GEN && a.nsrc_eh) {  // the old H without the last step's H source value
    T v[3] = {hx, hy, T(0)};
    soft_eh(a, step, 3, c, true, v);
    hx = v[0];
    hy = v[1];
  }
  const T im = a.mim[a.mat[c]];
  const T dxy = ez_yp - ez;  // Hx, J=y, +1
  const T dyx = ez_xp - ez;  // Hy, J=x, -1
  T d[3] = {T(0), T(0), T(0)};
  d[0] = -(a.cf[1][1][j] * dxy);
  d[1] = a.cf[1][0][i] * dyx;
  if (a.nsrc) add_sources(a, step, 2, c, d);
  const int sxi = slab_index(i, a.slab_a[0], a.slab_b[0]);
  if (sxi >= 0) {
    const long o = slab_off(a, 0, sxi, i, j, 0);
    const T p = a.psi[1][1][0][o];
    d[1] = fmadd(-a.cg[1][0][i], p, d[1]);
    a.psio[1][1][0][o] = fmadd(a.cb[1][0][i], p, -(a.cc[1][0][i] * dyx));
  }
  const int syi = slab_index(j, a.slab_a[1], a.slab_b[1]);
  if (syi >= 0) {
    const long o = slab_off(a, 1, syi, i, j, 0);
    const T p = a.psi[1][0][1][o];
    d[0] = fmadd(-a.cg[1][1][j], p, d[0]);
    a.psio[1][0][1][o] = fmadd(a.cb[1][1][j], p, a.cc[1][1][j] * dxy);
  }
  if (a.tfsf && (i == a.box[0] || i == a.box[1] || j == a.box[2] || j == a.box[3])) {
    const bool aF[3] = {true, true, false}, aC[3] = {false, false, true};
    tfsf_cell<T, 2>(a, a.incE + (long)s * a.auxL, 1, i, j, 0, aF, aC, (const T*)nullptr, d);
  }
  if (GEN && a.nplane) plane_cells<T, 2>(a, step, 1, i, j, 0, 3, 4, d);
  T h[3] = {fmadd(d[0], im, hx), fmadd(d[1], im, hy), T(0)};
  if (GEN && im < T(0)) {  // magnetic poles (the general instance only)
    const long bq = mbox_index(a, i, j, 0);
    h[0] = mag_component<T, 0>(a, c, bq, d[0], hx);
    h[1] = mag_component<T, 1>(a, c, bq, d[1], hy);
  }
  if (GEN && a.nsrc_eh) soft_eh(a, step, 3, c, false, h);
  a.Ho[0][c] = h[0];
  a.Ho[1][c] = h[1];
  if (GEN && a.bc_any) bc_ghost_write(a, 1, 2, i, j, 0, c, h);
  if (a.nprobe) {
    const T e[3] = {T(0), T(0), ez};
    dft_probe(a, step, c, i, j, 0, e, h);
  }
}

// ---------------------------------------------------------------- 2-D TE
// Ex, Ey, Hz (FieldComponents<TE>, include/FDTDConstants.hpp:852); CPML along
// x and y; in place.  The curl and CPML terms are those of e_cell_3d /
// h_cell_3d with the z derivatives and the TM components absent:
// Dx += dt F_y/dx (Hz - Hz[j-1]), Dy -= dt F_x/dx (Hz - Hz[i-1]),
// Bz -= dt (F_x/dx (Ey[i+1] - Ey) - F_y/dx (Ex[j+1] - Ex)).
template <typename T, bool GEN = false>
__global__ __launch_bounds__(256) void e_pass_te(const StepArgs<T>* __restrict__ ap, int, int s) {
#pragma clang fp contract(off)
  const StepArgs<T>& a = *ap;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i < 1 || i >= a.nx - 1 || j >= a.ny - 1) return;
  const FDTD_G T* __restrict__ Hz = a.H[2];
  const long step = a.step0 + s;
  const long c = (long)i + a.sy * j;
  const T hz = Hz[c], hz_ym = Hz[c - a.sy], hz_xm = Hz[c - 1];
  T ex = a.E[0][c], ey = a.E[1][c];
  if (GEN && a.nsrc_eh) {  // the old E without the last step's E source value
    T v[3] = {ex, ey, T(0)};
    soft_eh(a, step, 1, c, true, v);
    ex = v[0];
    ey = v[1];
  }
  const int mid = a.mat[c];
  const T ie = cell_inv_eps(a, a.mie, mid, i, j, 0);
  const T dxy = hz - hz_ym;  // Ex, J=y, +1
  const T dyx = hz - hz_xm;  // Ey, J=x, -1
  T d[3] = {T(0), T(0), T(0)};
  d[0] = a.cf[0][1][j] * dxy;
  d[1] = -(a.cf[0][0][i] * dyx);
  if (a.nsrc) add_sources(a, step, 0, c, d);
  const int sxi = slab_index(i, a.slab_a[0], a.slab_b[0]);
  if (sxi >= 0) {
    const long o = slab_off(a, 0, sxi, i, j, 0);
    const T p = a.psi[0][1][0][o];
    d[1] = fmadd(a.cg[0][0][i], p, d[1]);
    a.psio[0][1][0][o] = fmadd(a.cb[0][0][i], p, -(a.cc[0][0][i] * dyx));
  }
  const int syi = slab_index(j, a.slab_a[1], a.slab_b[1]);
  if (syi >= 0) {
    const long o = slab_off(a, 1, syi, i, j, 0);
    const T p = a.psi[0][0][1][o];
    d[0] = fmadd(a.cg[0][1][j], p, d[0]);
    a.psio[0][0][1][o] = fmadd(a.cb[0][1][j], p, a.cc[0][1][j] * dxy);
  }
  if (a.tfsf && (i == a.box[0] || i == a.box[1] || j == a.box[2] || j == a.box[3])) {
    const bool aF[3] = {true, true, false}, aC[3] = {false, false, true};
    tfsf_cell<T, 2>(a, a.incH + (long)s * a.auxL, 0, i, j, 0, aF, aC, (const T*)nullptr, d);
  }
  if (GEN && a.nplane) plane_cells<T, 2>(a, step, 0, i, j, 0, 3, 4, d);
  T e[3] = {ex, ey, T(0)};
  if (ie > T(0)) {
    e[0] = fmadd(d[0], ie, ex);
    e[1] = fmadd(d[1], ie, ey);
  } else if (ie < T(0) && a.Dd[0]) {
    const MatDev<T>& m = a.mats[mid];
    const long bq = dbox_index(a, i, j, 0);
    if (m.fluid) {
      e[0] = fluid_component<T, 0>(a, m, bq, d[0], true);
      e[1] = fluid_component<T, 1>(a, m, bq, d[1], true);
    } else {
      e[0] = ade_component<T, 0>(a, m, bq, d[0], ex, true);
      e[1] = ade_component<T, 1>(a, m, bq, d[1], ey, true);
    }
  }
  if (GEN && a.nsrc_eh) soft_eh(a, step, 1, c, false, e);
  a.Eo[0][c] = e[0];
  a.Eo[1][c] = e[1];
  if (GEN && a.bc_any) bc_ghost_write(a, 0, 2, i, j, 0, c, e);
}

template <typename T, bool GEN = false>
__global__ __launch_bounds__(256) void h_pass_te(const StepArgs<T>* __restrict__ ap, int, int s) {
#pragma clang fp contract(off)
  const StepArgs<T>& a = *ap;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i < 1 || i >= a.nx - 1 || j >= a.ny - 1) return;
  const FDTD_G T* __restrict__ Ex = a.Eo[0];
  const FDTD_G T* __restrict__ Ey = a.Eo[1];
  const long step = a.step0 + s;
  const long c = (long)i + a.sy * j;
  const T ex = Ex[c], ey = Ey[c], ex_yp = Ex[c + a.sy], ey_xp = Ey[c + 1];
  T hz = a.H[2][c];
  if (GEN && a.nsrc_eh) {  // the old H without the last step's H source value
    T v[3] = {T(0), T(0), hz};
    soft_eh(a, step, 3, c, true, v);
    hz = v[2];
  }
  const T im = a.mim[a.mat[c]];
  const T dzx = ey_xp - ey;  // Hz, J=x, +1
  const T dzy = ex_yp - ex;  // Hz, J=y, -1
  T d[3] = {T(0), T(0), T(0)};
  d[2] = fmadd(a.cf[1][1][j], dzy, -(a.cf[1][0][i] * dzx));
  if (a.nsrc) add_sources(a, step, 2, c, d);
  const int sxi = slab_index(i, a.slab_a[0], a.slab_b[0]);
  if (sxi >= 0) {
    const long o = slab_off(a, 0, sxi, i, j, 0);
    const T p = a.psi[1][2][0][o];
    d[2] = fmadd(-a.cg[1][0][i], p, d[2]);
    a.psio[1][2][0][o] = fmadd(a.cb[1][0][i], p, a.cc[1][0][i] * dzx);
  }
  const int syi = slab_index(j, a.slab_a[1], a.slab_b[1]);
  if (syi >= 0) {
    const long o = slab_off(a, 1, syi, i, j, 0);
    const T p = a.psi[1][2][1][o];
    d[2] = fmadd(-a.cg[1][1][j], p, d[2]);
    a.psio[1][2][1][o] = fmadd(a.cb[1][1][j], p, -(a.cc[1][1][j] * dzy));
  }
  if (a.tfsf && (i == a.box[0] || i == a.box[1] || j == a.box[2] || j == a.box[3])) {
    const bool aF[3] = {false, false, true}, aC[3] = {true, true, false};
    tfsf_cell<T, 2>(a, a.incE + (long)s * a.auxL, 1, i, j, 0, aF, aC, (const T*)nullptr, d);
  }
  if (GEN && a.nplane) plane_cells<T, 2>(a, step, 1, i, j, 0, 4, 3, d);
  T h[3] = {T(0), T(0), fmadd(d[2], im, hz)};
  if (GEN && im < T(0)) h[2] = mag_component<T, 2>(a, c, mbox_index(a, i, j, 0), d[2], hz);  // magnetic poles
  if (GEN && a.nsrc_eh) soft_eh(a, step, 3, c, false, h);
  a.Ho[2][c] = h[2];
  if (GEN && a.bc_any) bc_ghost_write(a, 1, 2, i, j, 0, c, h);
  if (a.nprobe) {
    const T e[3] = {ex, ey, T(0)};
    dft_probe(a, step, c, i, j, 0, e, h);
  }
}

// ---------------------------------------------------------------- 1-D TEM
// Ez, Hy along x (FieldComponents<TEM>, include/FDTDConstants.hpp:868-874);
// CPML along x; in place; one thread per cell.  The x terms of the TM passes:
// Dz += dt F_x/dx (Hy - Hy[i-1]), By += dt F_x/dx (Ez[i+1] - Ez).
template <typename T, bool GEN = false>
__global__ __launch_bounds__(256) void e_pass_tem(const StepArgs<T>* __restrict__ ap, int, int s) {
#pragma clang fp contract(off)
  const StepArgs<T>& a = *ap;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i >= a.nx - 1) return;
  const long step = a.step0 + s;
  const long c = i;
  const T hy = a.H[1][c], hy_xm = a.H[1][c - 1];
  T ez = a.E[2][c];
  if (GEN && a.nsrc_eh) {  // the old Ez without the last step's E source value
    T v[3] = {T(0), T(0), ez};
    soft_eh(a, step, 1, c, true, v);
    ez = v[2];
  }
  const int mid = a.mat[c];
  const T ie = cell_inv_eps(a, a.mie, mid, i, 0, 0);
  const T dzx = hy - hy_xm;  // J=x, +1
  T d[3] = {T(0), T(0), a.cf[0][0][i] * dzx};
  if (a.nsrc) add_sources(a, step, 0, c, d);
  const int sxi = slab_index(i, a.slab_a[0], a.slab_b[0]);
  if (sxi >= 0) {
    const long o = slab_off(a, 0, sxi, i, 0, 0);
    const T p = a.psi[0][2][0][o];
    d[2] = fmadd(a.cg[0][0][i], p, d[2]);
    a.psio[0][2][0][o] = fmadd(a.cb[0][0][i], p, a.cc[0][0][i] * dzx);
  }
  if (a.tfsf && (i == a.box[0] || i == a.box[1])) {
    const bool aF[3] = {false, false, true}, aC[3] = {false, true, false};
    tfsf_cell<T, 1>(a, a.incH + (long)s * a.auxL, 0, i, 0, 0, aF, aC, (const T*)nullptr, d);
  }
  if (GEN && a.nplane) plane_cells<T, 1>(a, step, 0, i, 0, 0, 4, 2, d);
  T ezn = ez;
  if (ie > T(0)) {
    ezn = fmadd(d[2], ie, ez);
  } else if (ie < T(0) && a.Dd[2]) {
    const MatDev<T>& m = a.mats[mid];
    if (m.fluid)
      ezn = fluid_component<T, 2>(a, m, dbox_index(a, i, 0, 0), d[2], true);
    else
      ezn = ade_component<T, 2>(a, m, dbox_index(a, i, 0, 0), d[2], ez, true);
  }
  T e[3] = {T(0), T(0), ezn};
  if (GEN && a.nsrc_eh) soft_eh(a, step, 1, c, false, e);
  a.Eo[2][c] = e[2];
  if (GEN && a.bc_any) bc_ghost_write(a, 0, 1, i, 0, 0, c, e);
}

template <typename T, bool GEN = false>
__global__ __launch_bounds__(256) void h_pass_tem(const StepArgs<T>* __restrict__ ap, int, int s) {
#pragma clang fp contract(off)
  const StepArgs<T>& a = *ap;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i >= a.nx - 1) return;
  const long step = a.step0 + s;
  const long c = i;
  const T ez = a.Eo[2][c], ez_xp = a.Eo[2][c + 1];
  T hy = a.H[1][c];
  if (GEN && a.nsrc_eh) {  // the old H without the last step's H source value
    T v[3] = {T(0), hy, T(0)};
    soft_eh(a, step, 3, c, true, v);
    hy = v[1];
  }
  const T im = a.mim[a.mat[c]];
  const T dyx = ez_xp - ez;  // Hy, J=x, -1
  T d[3] = {T(0), a.cf[1][0][i] * dyx, T(0)};
  if (a.nsrc) add_sources(a, step, 2, c, d);
  const int sxi = slab_index(i, a.slab_a[0], a.slab_b[0]);
  if (sxi >= 0) {
    const long o = slab_off(a, 0, sxi, i, 0, 0);
    const T p = a.psi[1][1][0][o];
    d[1] = fmadd(-a.cg[1][0][i], p, d[1]);
    a.psio[1][1][0][o] = fmadd(a.cb[1][0][i], p, -(a.cc[1][0][i] * dyx));
  }
  if (a.tfsf && (i == a.box[0] || i == a.box[1])) {
    const bool aF[3] = {false, true, false}, aC[3] = {false, false, true};
    tfsf_cell<T, 1>(a, a.incE + (long)s * a.auxL, 1, i, 0, 0, aF, aC, (const T*)nullptr, d);
  }
  if (GEN && a.nplane) plane_cells<T, 1>(a, step, 1, i, 0, 0, 2, 4, d);
  T h[3] = {T(0), fmadd(d[1], im, hy), T(0)};
  if (GEN && im < T(0)) h[1] = mag_component<T, 1>(a, c, mbox_index(a, i, 0, 0), d[1], hy);  // magnetic poles
  if (GEN && a.nsrc_eh) soft_eh(a, step, 3, c, false, h);
  a.Ho[1][c] = h[1];
  if (GEN && a.bc_any) bc_ghost_write(a, 1, 1, i, 0, 0, c, h);
  if (a.nprobe) {
    const T ev[3] = {T(0), T(0), ez};
    dft_probe(a, step, c, i, 0, 0, ev, h);
  }
}

// ---------------------------------------------------------------- 2-D TM, time-blocked
// One workgroup per C x C core of the grid advances it TB steps per launch:
// it loads the core and a halo of TB cells on every side (the state at the
// launch's first step), keeps one cell per thread in registers, passes the
// neighbour values of each half step through LDS, and writes the core into
// the other copy of the state.  Each step's E update consumes one halo layer
// on the low side (it reads H at -x, -y) and the H update one on the high
// side (E at +x, +y), so after TB steps the core is exact; each step updates
// only the cells the core still depends on.  The redundant halo work buys TB
// steps per launch on a grid that fits in the caches and whose one-pass steps
// are bound by launch latency.
// Grids with dispersive cells, boundary conditions, E / H sources, single
// planes, complex fields or DFT monitors take the two passes.  The cell
// arithmetic is that of e_pass_tm / h_pass_tm, operation for operation.
constexpr int kTmMaxRegion = 32;  // C + 2 TB <= 32: one thread per region cell, 1024 threads
constexpr int kTmMaxSteps = 8;
constexpr int kTmRows = 3072;  // incident-row values (both fields, TB steps) staged in LDS at most
template <typename T, int TB>
__global__ __launch_bounds__(1024) void tm_block(const StepArgs<T>* __restrict__ ap, int core, int s0, int nst,
                                                 int ntx) {
#pragma clang fp contract(off)
  const StepArgs<T>& a = *ap;
  __shared__ T sx[kTmMaxRegion * kTmMaxRegion];  // Hx
  __shared__ T sy[kTmMaxRegion * kTmMaxRegion];  // Hy
  __shared__ T se[kTmMaxRegion * kTmMaxRegion];  // Ez
  // the launch's source values, [source][step], their lengths and components
  __shared__ T sval[kMaxSrc][TB];
  __shared__ int slen[kMaxSrc], scomp[kMaxSrc];
  // the launch's incident rows: Hy for the D faces, Ez for the B faces
  __shared__ double rowH[kTmRows / 2], rowE[kTmRows / 2];
  const int R = core + 2 * TB;
  const int t = threadIdx.x;
  const bool live = t < R * R;
  const int rx = live ? t % R : 0, ry = live ? t / R : 0;
  const int bx = blockIdx.x % ntx, by = blockIdx.x / ntx;
  const int i = bx * core - TB + rx, j = by * core - TB + ry;
  const bool ingrid = live && i >= 0 && j >= 0 && i < a.nx && j < a.ny;
  const bool inner = ingrid && i >= 1 && j >= 1 && i < a.nx - 1 && j < a.ny - 1;
  const long c = (long)i + a.sy * j;
  // stage the step-dependent tables of the launch (no global loads in the step loop)
  const long stepA = a.step0 + s0;
  for (int q = t; q < a.nsrc * TB; q += blockDim.x) {
    const int sidx = q / TB, k = q % TB;
    const SrcDev& sd = a.src[sidx];
    sval[sidx][k] = (k < nst && stepA + k < sd.n) ? (T)sd.table[stepA + k] : T(0);
    if (k == 0) {
      slen[sidx] = sd.n;
      scomp[sidx] = sd.comp;
    }
  }
  const bool rows_lds = a.tfsf && a.auxL * nst <= kTmRows / 2;
  if (rows_lds)
    for (int q = t; q < a.auxL * nst; q += blockDim.x) {
      rowH[q] = a.incH[(long)s0 * a.auxL + q];
      rowE[q] = a.incE[(long)s0 * a.auxL + q];
    }
  T ez = T(0), hx = T(0), hy = T(0);
  if (ingrid) {
    ez = a.E[2][c];
    hx = a.H[0][c];
    hy = a.H[1][c];
  }
  // the cell's coefficients and CPML state
  T ie = T(0), im = T(0), fEx = T(0), fEy = T(0), fHx = T(0), fHy = T(0);
  const int sxi = inner ? slab_index(i, a.slab_a[0], a.slab_b[0]) : -1;
  const int syi = inner ? slab_index(j, a.slab_a[1], a.slab_b[1]) : -1;
  long ox = 0, oy = 0;
  T pEx = T(0), pEy = T(0), pHx = T(0), pHy = T(0);  // psi of Dz (x, y slabs), of By (x slab), Bx (y slab)
  T gEx = T(0), bEx = T(0), cEx = T(0), gEy = T(0), bEy = T(0), cEy = T(0);
  T gHx = T(0), bHx = T(0), cHx = T(0), gHy = T(0), bHy = T(0), cHy = T(0);
  unsigned smD = 0, smB = 0;  // the D / B sources of this cell (bit per source)
  if (inner) {
    const int mid = a.mat[c];
    ie = cell_inv_eps(a, a.mie, mid, i, j, 0);
    im = a.mim[mid];
    fEx = a.cf[0][0][i];
    fEy = a.cf[0][1][j];
    fHx = a.cf[1][0][i];
    fHy = a.cf[1][1][j];
    if (sxi >= 0) {
      ox = slab_off(a, 0, sxi, i, j, 0);
      pEx = a.psi[0][2][0][ox];
      pHx = a.psi[1][1][0][ox];
      gEx = a.cg[0][0][i];
      bEx = a.cb[0][0][i];
      cEx = a.cc[0][0][i];
      gHx = a.cg[1][0][i];
      bHx = a.cb[1][0][i];
      cHx = a.cc[1][0][i];
    }
    if (syi >= 0) {
      oy = slab_off(a, 1, syi, i, j, 0);
      pEy = a.psi[0][2][1][oy];
      pHy = a.psi[1][0][1][oy];
      gEy = a.cg[0][1][j];
      bEy = a.cb[0][1][j];
      cEy = a.cc[0][1][j];
      gHy = a.cg[1][1][j];
      bHy = a.cb[1][1][j];
      cHy = a.cc[1][1][j];
    }
    for (int q = 0; q < a.nsrc; q++)
      if (a.src[q].cell == c) {
        if (a.src[q].field == 0) smD |= 1u << q;
        if (a.src[q].field == 2) smB |= 1u << q;
      }
  }
  const bool owner = inner && rx >= TB && ry >= TB && rx < TB + core && ry < TB + core;
  // the probes this thread records: bit p of mEz / mHx / mHy
  unsigned mEz = 0, mHx = 0, mHy = 0;
  if (owner)
    for (int p = 0; p < a.nprobe; p++)
      if (a.probe[p].cell == c) {
        const int f = a.probe[p].field, cp = a.probe[p].comp;
        if (f == 1 && cp == 2) mEz |= 1u << p;
        if (f == 3 && cp == 0) mHx |= 1u << p;
        if (f == 3 && cp == 1) mHy |= 1u << p;
      }
  const bool tfsf = inner && a.tfsf && (i == a.box[0] || i == a.box[1] || j == a.box[2] || j == a.box[3]);
  __syncthreads();  // the staged tables
  for (int q = 0; q < nst; q++) {
    const int s = s0 + q;
    const long step = a.step0 + s;
    // the cells whose new values the core still depends on: E on [lo, hi],
    // H on [lo, hi - 1] along each axis (the dependency cone of the core
    // after the launch's remaining steps; lo >= 1 and hi <= R - 1, so the
    // neighbours read are in the region)
    const int lo = TB - (nst - 1 - q), hi = TB + core + (nst - 1 - q);
    // E half step: Ez from Hx, Hy and their -y, -x neighbours
    if (live) {
      sx[t] = hx;
      sy[t] = hy;
    }
    s_bar_lds();
    if (inner && rx >= lo && ry >= lo && rx <= hi && ry <= hi) {
      const T dzx = hy - sy[t - 1];
      const T dzy = hx - sx[t - R];
      T d[3] = {T(0), T(0), T(0)};
      d[2] = fmadd(fEx, dzx, -(fEy * dzy));
      for (unsigned m = smD; m; m &= m - 1) {  // add_sources, from the staged values
        const int sq = __builtin_ctz(m);
        if (step < slen[sq]) d[2] += sval[sq][q];
      }
      if (sxi >= 0) {
        d[2] = fmadd(gEx, pEx, d[2]);
        pEx = fmadd(bEx, pEx, cEx * dzx);
      }
      if (syi >= 0) {
        d[2] = fmadd(gEy, pEy, d[2]);
        pEy = fmadd(bEy, pEy, -(cEy * dzy));
      }
      if (tfsf) {
        const bool aF[3] = {false, false, true}, aC[3] = {true, true, false};
        if (rows_lds)
          tfsf_cell<T, 2, false, const double*>(a, rowH + q * a.auxL, 0, i, j, 0, aF, aC, (const T*)nullptr, d);
        else
          tfsf_cell<T, 2>(a, a.incH + (long)s * a.auxL, 0, i, j, 0, aF, aC, (const T*)nullptr, d);
      }
      if (ie > T(0)) ez = fmadd(d[2], ie, ez);
    }
    if (live) se[t] = ez;
    s_bar_lds();
    // H half step: Hx, Hy from Ez and its +y, +x neighbours (the next step
    // writes sx, sy only after every thread has passed this barrier, that
    // is after every E update read them; se likewise after the next one)
    if (inner && rx >= lo && ry >= lo && rx < hi && ry < hi) {
      const T dxy = se[t + R] - ez;
      const T dyx = se[t + 1] - ez;
      T d[3] = {T(0), T(0), T(0)};
      d[0] = -(fHy * dxy);
      d[1] = fHx * dyx;
      for (unsigned m = smB; m; m &= m - 1) {
        const int sq = __builtin_ctz(m);
        if (step < slen[sq]) {
          if (scomp[sq] == 0)
            d[0] += sval[sq][q];
          else
            d[1] += sval[sq][q];
        }
      }
      if (sxi >= 0) {
        d[1] = fmadd(-gHx, pHx, d[1]);
        pHx = fmadd(bHx, pHx, -(cHx * dyx));
      }
      if (syi >= 0) {
        d[0] = fmadd(-gHy, pHy, d[0]);
        pHy = fmadd(bHy, pHy, cHy * dxy);
      }
      if (tfsf) {
        const bool aF[3] = {true, true, false}, aC[3] = {false, false, true};
        if (rows_lds)
          tfsf_cell<T, 2, false, const double*>(a, rowE + q * a.auxL, 1, i, j, 0, aF, aC, (const T*)nullptr, d);
        else
          tfsf_cell<T, 2>(a, a.incE + (long)s * a.auxL, 1, i, j, 0, aF, aC, (const T*)nullptr, d);
      }
      hx = fmadd(d[0], im, hx);
      hy = fmadd(d[1], im, hy);
    }
    for (unsigned m = mEz; m; m &= m - 1) a.trace[step * a.nprobe + __builtin_ctz(m)] = (double)ez;
    for (unsigned m = mHx; m; m &= m - 1) a.trace[step * a.nprobe + __builtin_ctz(m)] = (double)hx;
    for (unsigned m = mHy; m; m &= m - 1) a.trace[step * a.nprobe + __builtin_ctz(m)] = (double)hy;
  }
  // the core, into the other copy of the state (the outer layer unchanged)
  if (ingrid && rx >= TB && ry >= TB && rx < TB + core && ry < TB + core) {
    a.Eo[2][c] = ez;
    a.Ho[0][c] = hx;
    a.Ho[1][c] = hy;
    if (sxi >= 0) {
      a.psio[0][2][0][ox] = pEx;
      a.psio[1][1][0][ox] = pHx;
    }
    if (syi >= 0) {
      a.psio[0][2][1][oy] = pEy;
      a.psio[1][0][1][oy] = pHy;
    }
  }
}

// ---------------------------------------------------------------- Aux1D line
// One workgroup advances the 1-D TEM incident line `nsteps` times
// (Aux1D::time_step, include/Source.hpp:413-435) and writes, per step, the
// line's Hy before the advance and its Ez after it.  Always FP64, in the
// operation order of the reference's update (no contraction).
// Thread t owns the CPT consecutive cells [t CPT, t CPT + CPT): their Dz, By,
// Ez, Hy and curl coefficients stay in registers for the whole launch; only
// the first Ez and the last Hy of each thread cross to its neighbours, through
// LDS, with barriers that wait for LDS alone (the row stores stay in flight).
// The CPML cells (the first npml+2 and the last npml+3) keep psi and their
// B, C, G coefficients in an LDS table; the signal is staged in LDS per block
// of kAuxSigBlock steps.  Cells past L hold zero and are never updated.
struct AuxArgs {
  int L, npml;
  double dt, dx, eps0, mu0;
  double* Dz;
  double* Ez;
  double* By;
  double* Hy;
  double* psiE;
  double* psiH;
  const double* pE;  // [L][4] B,C,F,G
  const double* pH;
  const double* signal;
  int nsignal;
  long step0;
  int nsteps;
  double* outE;  // [nsteps][L]
  double* outH;
};

constexpr int kAuxThreads = 1024;
constexpr int kAuxPmlMax = 64;  // CPML cells of a line: 2 npml + 5 (npml <= 29)
constexpr int kAuxSigBlock = 256;

template <int CPT>
__global__ __launch_bounds__(kAuxThreads) void aux_line(AuxArgs a) {
#pragma clang fp contract(off)
  __shared__ double e_first[kAuxThreads], h_last[kAuxThreads];
  __shared__ double pml_psi[2][kAuxPmlMax], pml_c[2][3][kAuxPmlMax];  // [E/H][B, C, G]
  __shared__ double sig[kAuxSigBlock];
  const int L = a.L, t = threadIdx.x, nt = blockDim.x;
  const double dt = a.dt, dx = a.dx;
  const int src = a.npml + 5, plo = a.npml + 2, phi = L - (a.npml + 3);
  // table slot of a CPML cell, -1 for the others
  auto pslot = [&](int n) { return n < plo ? n : (n >= phi && n < L ? plo + (n - phi) : -1); };
  double D[CPT], B[CPT], E[CPT], H[CPT], fE[CPT], fH[CPT];
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const int n = t * CPT + q;
    const bool in = n < L;
    D[q] = in ? a.Dz[n] : 0.0;
    B[q] = in ? a.By[n] : 0.0;
    E[q] = in ? a.Ez[n] : 0.0;
    H[q] = in ? a.Hy[n] : 0.0;
    fE[q] = in ? a.pE[4 * n + 2] : 0.0;
    fH[q] = in ? a.pH[4 * n + 2] : 0.0;
    const int ps = in ? pslot(n) : -1;
    if (ps >= 0) {
      pml_psi[0][ps] = a.psiE[n];
      pml_psi[1][ps] = a.psiH[n];
      pml_c[0][0][ps] = a.pE[4 * n + 0];
      pml_c[0][1][ps] = a.pE[4 * n + 1];
      pml_c[0][2][ps] = a.pE[4 * n + 3];
      pml_c[1][0][ps] = a.pH[4 * n + 0];
      pml_c[1][1][ps] = a.pH[4 * n + 1];
      pml_c[1][2][ps] = a.pH[4 * n + 3];
    }
  }
  h_last[t] = H[CPT - 1];
  // does this thread own a CPML cell or the source cell (wave-uniform tests keep
  // the other waves' loops free of the table code)
  const bool has_pml = t * CPT < plo || t * CPT + CPT - 1 >= phi;
  const bool has_src = src >= t * CPT && src < t * CPT + CPT;
  for (int s = 0; s < a.nsteps; s++) {
    if (s % kAuxSigBlock == 0) {
      __syncthreads();
      for (int x = t; x < kAuxSigBlock; x += nt) {
        const long gs = a.step0 + s + x;
        sig[x] = gs < a.nsignal ? a.signal[gs] : 0.0;
      }
      __syncthreads();
    }
    FDTD_G double* oH = (FDTD_G double*)a.outH + (long)s * L;
    FDTD_G double* oE = (FDTD_G double*)a.outE + (long)s * L;
#pragma unroll
    for (int q = 0; q < CPT; q++)
      if (t * CPT + q < L) oH[t * CPT + q] = H[q];
    // E: Dz += dt F/dx (Hy - Hy[n-1]); CPML; Ez = Dz/eps0 (+ the signal at the source cell)
    double hm = t > 0 ? h_last[t - 1] : 0.0;
#pragma unroll
    for (int q = 0; q < CPT; q++) {
      const int n = t * CPT + q;
      const double diff = H[q] - hm;
      hm = H[q];
      if (n < L) {
        D[q] = D[q] + dt * (fE[q] / dx * diff);
        if (has_pml) {
          const int ps = pslot(n);
          if (ps >= 0) {
            const double p = pml_psi[0][ps];
            D[q] += dt * p * pml_c[0][2][ps];
            pml_psi[0][ps] = pml_c[0][0][ps] * p + pml_c[0][1][ps] / dx * diff;
          }
        }
        E[q] = D[q] / a.eps0;
        if (has_src && n == src) E[q] += sig[s % kAuxSigBlock];
      }
    }
    e_first[t] = E[0];
    s_bar_lds();
    // B: By += dt F/dx (Ez[n+1] - Ez); CPML; Hy = By/mu0
    const double enext = t + 1 < nt ? e_first[t + 1] : 0.0;
#pragma unroll
    for (int q = 0; q < CPT; q++) {
      const int n = t * CPT + q;
      const double ep = q + 1 < CPT ? E[q + 1] : enext;
      if (n < L) {
        const double diff = ep - E[q];
        B[q] = B[q] + dt * (fH[q] / dx * diff);
        if (has_pml) {
          const int ps = pslot(n);
          if (ps >= 0) {
            const double p = pml_psi[1][ps];
            B[q] -= dt * p * pml_c[1][2][ps];
            pml_psi[1][ps] = pml_c[1][0][ps] * p - pml_c[1][1][ps] / dx * diff;
          }
        }
        oE[n] = E[q];
        H[q] = B[q] / a.mu0;
      }
    }
    h_last[t] = H[CPT - 1];
    s_bar_lds();
  }
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const int n = t * CPT + q;
    if (n < L) {
      a.Dz[n] = D[q];
      a.By[n] = B[q];
      a.Ez[n] = E[q];
      a.Hy[n] = H[q];
      const int ps = pslot(n);
      if (ps >= 0) {
        a.psiE[n] = pml_psi[0][ps];
        a.psiH[n] = pml_psi[1][ps];
      }
    }
  }
}

}  // namespace fdtdk
