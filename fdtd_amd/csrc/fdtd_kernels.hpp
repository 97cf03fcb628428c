// Device side of the FDTD hot path for gfx950.
//
// Two passes per time step, each one HBM stream over the grid:
//   e_pass: D curl + soft sources + CPML(E) + TFSF updateD + E material/ADE
//   h_pass: B curl + sources + CPML(H) + TFSF updateB + H material, then the
//           running DFT and the probes of this step
// Fields are structure-of-arrays (one array per component, x fastest), so a
// 64-lane wavefront reads 64 consecutive cells of one component.  Each thread
// owns one (i,j) column and marches a chunk of z planes, keeping the plane it
// just read in registers, so the z-neighbour of the curl costs no reload.
// E and H are stored directly (E = D/(eps_r eps0) folded into the update);
// D, P and J exist only inside the dispersive box.
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

namespace fdtdk {

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
};

struct SrcDev {
  long cell;
  int field, comp;
  const double* table;
  int n;
};

template <typename T>
struct DftDev {
  int r[6];
  int mask, nfreq, ncomp;
  long ncells;
  T* acc;               // [ncomp][nfreq][2][ncells]
  const double* phE;    // [nphase][nfreq][2]
  const double* phH;
  int nphase;
};

struct ProbeDev {
  long cell;
  int field, comp;
};

template <typename T>
struct StepArgs {
  T* E[3];
  T* H[3];
  const uint8_t* mat;
  const MatDev<T>* mats;
  const T* mie;  // 1/(eps_r eps0) per material id, negated for dispersive materials
  const T* mim;  // 1/(mu_r mu0) per material id
  // CPML: psi[ft][I][J] slab arrays and per-axis coefficient profiles
  T* psi[2][3][3];
  const T* cf[2][3];  // dt F / dx
  const T* cg[2][3];  // dt G
  const T* cb[2][3];  // B
  const T* cc[2][3];  // C / dx
  int slab_a[3], slab_b[3], slab_S[3];
  int nx, ny, nz;
  long sy, sz;
  // dispersive box (dense): [x0,x1)x[y0,y1)x[z0,z1)
  int dbox[6];
  T* Dd[3];
  T* Pd[kMaxPoles][3];
  T* Jd[kMaxPoles][3];
  T dt;
  // soft sources
  int nsrc;
  SrcDev src[kMaxSrc];
  // TFSF
  int tfsf;
  int box[6];
  double kvec[3], pol[3], polH[3], origin[3], dx;
  int auxL;
  const double* incE;  // aux Ez after this step's aux advance
  const double* incH;  // aux Hy before it
  // running DFT
  int ndft;
  DftDev<T> dft[kMaxDft];
  // probes
  int nprobe;
  ProbeDev probe[kMaxProbe];
  double* trace;  // [step][nprobe]
  long step;
};

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
__device__ __forceinline__ double interp_row(const double* row, int L, double loc, double dx, bool h) {
  double nd = (double)L / 2.0 + loc / dx;
  if (h) nd -= 0.5;
  long n = (long)nd;
  if (n < 0) n = 0;
  if (n > L - 2) n = L - 2;
  double frac = nd - (double)n;
  return (1.0 - frac) * row[n] + frac * row[n + 1];
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
template <typename T, int NDIM>
__device__ void tfsf_cell(const StepArgs<T>& a, int ft, int i, int j, int k, const bool actF[3], const bool actC[3],
                          T delta[3]) {
  const int ci[3] = {i, j, k};
  for (int J = 0; J < NDIM; J++) {
    int side = (ci[J] == a.box[2 * J]) ? 0 : ((ci[J] == a.box[2 * J + 1]) ? 1 : -1);
    if (side < 0) continue;
    const double nrm = side == 0 ? -1.0 : 1.0;
    for (int I = 0; I < 3; I++) {
      if (I == J || !actF[I]) continue;
      const int K = 3 - I - J;
      if (!actC[K]) continue;
      bool inside = true;
      for (int Tn = 0; Tn < NDIM; Tn++) {
        if (Tn == J) continue;
        bool half = ft == 0 ? (Tn != I) : (Tn == I);
        int lo = half ? a.box[2 * Tn] + 1 : a.box[2 * Tn];
        int hi = half ? a.box[2 * Tn + 1] : a.box[2 * Tn + 1] - 1;
        if (ci[Tn] < lo || ci[Tn] > hi) inside = false;
      }
      if (!inside) continue;
      double proj = 0.0;
      for (int dd = 0; dd < NDIM; dd++) {
        double off = ft == 0 ? (dd == K ? -0.5 : 0.0) : (dd != K ? -0.5 : 0.0);
        double loc = ((double)ci[dd] - a.origin[dd]) * a.dx + off * a.dx;
        proj += a.kvec[dd] * loc;
      }
      double inc = ft == 0 ? interp_row(a.incH, a.auxL, proj, a.dx, true) : interp_row(a.incE, a.auxL, proj, a.dx, false);
      double pv = ft == 0 ? a.polH[K] : a.pol[K];
      double f = (double)a.cf[ft][J][ci[J]];
      double term = lc3(I, J) * f * nrm * inc * pv;
      delta[I] += (T)(ft == 0 ? term : -term);
    }
  }
}

// ---------------------------------------------------------------- E pass
template <typename T>
__device__ __forceinline__ T pick3(T v0, T v1, T v2, int c) {
  return c == 0 ? v0 : (c == 1 ? v1 : v2);
}

template <typename T>
__device__ __forceinline__ void add_sources(const StepArgs<T>& a, int field, long c, T d[3]) {
  for (int s = 0; s < a.nsrc; s++) {
    const SrcDev& q = a.src[s];
    if (q.cell == c && q.field == field && a.step < q.n) {
      const T v = (T)q.table[a.step];
      if (q.comp == 0)
        d[0] += v;
      else if (q.comp == 1)
        d[1] += v;
      else
        d[2] += v;
    }
  }
}

// Dispersive cell, one component: D += delta; P_p += dt J_p;
// E = (D - sum_p P_p)/(eps_r eps0); J_p = cJ J_p + cE E + cP P_p
// (DrudeUpdate / LorentzUpdate implicit_update, DispersiveMaterials.hpp:710-735, 979-998).
template <typename T, int I>
__device__ __forceinline__ T ade_component(const StepArgs<T>& a, const MatDev<T>& m, long bq, T d) {
  const T D = a.Dd[I][bq] + d;
  a.Dd[I][bq] = D;
  T num = D;
  T Pv[kMaxPoles], Jv[kMaxPoles];
#pragma unroll
  for (int p = 0; p < kMaxPoles; p++) {
    if (p < m.np) {
      Jv[p] = a.Jd[p][I][bq];
      Pv[p] = a.Pd[p][I][bq] + a.dt * Jv[p];
      num -= Pv[p];
    }
  }
  const T Ev = num * m.inv_eps;
#pragma unroll
  for (int p = 0; p < kMaxPoles; p++) {
    if (p < m.np) {
      a.Pd[p][I][bq] = Pv[p];
      a.Jd[p][I][bq] = m.cJ[p] * Jv[p] + m.cE[p] * Ev + m.cP[p] * Pv[p];
    }
  }
  return Ev;
}

template <typename T>
__device__ __forceinline__ long dbox_index(const StepArgs<T>& a, int i, int j, int k) {
  return (long)(i - a.dbox[0]) +
         (long)(a.dbox[1] - a.dbox[0]) * ((long)(j - a.dbox[2]) + (long)(a.dbox[3] - a.dbox[2]) * (long)(k - a.dbox[4]));
}

// One thread per cell and z plane: every load of the cell is issued before
// any arithmetic, so a wave carries ~12 independent loads in flight and the
// stores go out at the end.  mie[m] < 0 marks a dispersive material.
template <typename T>
__global__ __launch_bounds__(256) void e_pass_3d(StepArgs<T> a, int kc) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k0 = 1 + blockIdx.z * kc;
  const int k1 = min(k0 + kc, a.nz - 1);
  if (i >= a.nx - 1 || j >= a.ny - 1) return;
  const T* __restrict__ Hx = a.H[0];
  const T* __restrict__ Hy = a.H[1];
  const T* __restrict__ Hz = a.H[2];
  T* __restrict__ Ex = a.E[0];
  T* __restrict__ Ey = a.E[1];
  T* __restrict__ Ez = a.E[2];
  const uint8_t* __restrict__ mat = a.mat;
  const T* __restrict__ mie = a.mie;
  const long sy = a.sy, sz = a.sz;
  const T fx = a.cf[0][0][i], fy = a.cf[0][1][j];
  const int sxi = slab_index(i, a.slab_a[0], a.slab_b[0]);
  const int syi = slab_index(j, a.slab_a[1], a.slab_b[1]);
  for (int k = k0; k < k1; k++) {
    const long c = (long)i + sy * (long)j + sz * (long)k;
    const T hx = Hx[c], hy = Hy[c], hz = Hz[c];
    const T hz_ym = Hz[c - sy], hz_xm = Hz[c - 1];
    const T hy_xm = Hy[c - 1], hy_zm = Hy[c - sz];
    const T hx_ym = Hx[c - sy], hx_zm = Hx[c - sz];
    const T ex = Ex[c], ey = Ey[c], ez = Ez[c];
    const T ie = mie[mat[c]];
    const T fz = a.cf[0][2][k];
    const T dxy = hz - hz_ym;  // Ex, J=y, +1
    const T dxz = hy - hy_zm;  // Ex, J=z, -1
    const T dyz = hx - hx_zm;  // Ey, J=z, +1
    const T dyx = hz - hz_xm;  // Ey, J=x, -1
    const T dzx = hy - hy_xm;  // Ez, J=x, +1
    const T dzy = hx - hx_ym;  // Ez, J=y, -1
    T d[3];
    d[0] = fy * dxy - fz * dxz;
    d[1] = fz * dyz - fx * dyx;
    d[2] = fx * dzx - fy * dzy;
    if (a.nsrc) add_sources(a, 0, c, d);
    if (sxi >= 0) {
      const long o = slab_off(a, 0, sxi, i, j, k);
      T* __restrict__ q1 = a.psi[0][1][0];
      T* __restrict__ q2 = a.psi[0][2][0];
      const T g = a.cg[0][0][i], b = a.cb[0][0][i], cc = a.cc[0][0][i];
      const T p1 = q1[o], p2 = q2[o];
      d[1] += g * p1;
      d[2] += g * p2;
      q1[o] = b * p1 - cc * dyx;
      q2[o] = b * p2 + cc * dzx;
    }
    if (syi >= 0) {
      const long o = slab_off(a, 1, syi, i, j, k);
      T* __restrict__ q0 = a.psi[0][0][1];
      T* __restrict__ q2 = a.psi[0][2][1];
      const T g = a.cg[0][1][j], b = a.cb[0][1][j], cc = a.cc[0][1][j];
      const T p0 = q0[o], p2 = q2[o];
      d[0] += g * p0;
      d[2] += g * p2;
      q0[o] = b * p0 + cc * dxy;
      q2[o] = b * p2 - cc * dzy;
    }
    const int szi = slab_index(k, a.slab_a[2], a.slab_b[2]);
    if (szi >= 0) {
      const long o = slab_off(a, 2, szi, i, j, k);
      T* __restrict__ q0 = a.psi[0][0][2];
      T* __restrict__ q1 = a.psi[0][1][2];
      const T g = a.cg[0][2][k], b = a.cb[0][2][k], cc = a.cc[0][2][k];
      const T p0 = q0[o], p1 = q1[o];
      d[0] += g * p0;
      d[1] += g * p1;
      q0[o] = b * p0 - cc * dxz;
      q1[o] = b * p1 + cc * dyz;
    }
    if (a.tfsf && (i == a.box[0] || i == a.box[1] || j == a.box[2] || j == a.box[3] || k == a.box[4] || k == a.box[5])) {
      const bool act[3] = {true, true, true};
      tfsf_cell<T, 3>(a, 0, i, j, k, act, act, d);
    }
    if (ie > T(0)) {
      Ex[c] = ex + d[0] * ie;
      Ey[c] = ey + d[1] * ie;
      Ez[c] = ez + d[2] * ie;
    } else if (ie < T(0) && a.Dd[0]) {
      const MatDev<T>& m = a.mats[mat[c]];
      const long bq = dbox_index(a, i, j, k);
      Ex[c] = ade_component<T, 0>(a, m, bq, d[0]);
      Ey[c] = ade_component<T, 1>(a, m, bq, d[1]);
      Ez[c] = ade_component<T, 2>(a, m, bq, d[2]);
    }
  }
}

// ---------------------------------------------------------------- H pass
template <typename T>
__device__ __forceinline__ void dft_probe(const StepArgs<T>& a, long c, int i, int j, int k, const T e[3], const T h[3]) {
  for (int q = 0; q < a.ndft; q++) {
    const DftDev<T>& d = a.dft[q];
    if (i < d.r[0] || i >= d.r[1] || j < d.r[2] || j >= d.r[3] || k < d.r[4] || k >= d.r[5]) continue;
    if (a.step >= d.nphase) continue;
    const long lq =
        (long)(i - d.r[0]) + (long)(d.r[1] - d.r[0]) * ((long)(j - d.r[2]) + (long)(d.r[3] - d.r[2]) * (k - d.r[4]));
    int slot = 0;
#pragma unroll
    for (int comp = 0; comp < 6; comp++) {
      if (!((d.mask >> comp) & 1)) continue;
      const T v = comp < 3 ? e[comp] : h[comp - 3];
      const double* ph = (comp < 3 ? d.phE : d.phH) + 2L * d.nfreq * a.step;
      T* acc = d.acc + (long)slot * d.nfreq * 2 * d.ncells + lq;
      for (int f = 0; f < d.nfreq; f++) {
        T* re = acc + (long)f * 2 * d.ncells;
        re[0] += v * (T)ph[2 * f];
        re[d.ncells] += v * (T)ph[2 * f + 1];
      }
      slot++;
    }
  }
  for (int p = 0; p < a.nprobe; p++)
    if (a.probe[p].cell == c) {
      const int comp = a.probe[p].comp;
      const T v = a.probe[p].field == 1 ? pick3(e[0], e[1], e[2], comp) : pick3(h[0], h[1], h[2], comp);
      a.trace[a.step * a.nprobe + p] = (double)v;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void h_pass_3d(StepArgs<T> a, int kc) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k0 = 1 + blockIdx.z * kc;
  const int k1 = min(k0 + kc, a.nz - 1);
  if (i >= a.nx - 1 || j >= a.ny - 1) return;
  const T* __restrict__ Ex = a.E[0];
  const T* __restrict__ Ey = a.E[1];
  const T* __restrict__ Ez = a.E[2];
  T* __restrict__ Hx = a.H[0];
  T* __restrict__ Hy = a.H[1];
  T* __restrict__ Hz = a.H[2];
  const uint8_t* __restrict__ mat = a.mat;
  const T* __restrict__ mim = a.mim;
  const long sy = a.sy, sz = a.sz;
  const T fx = a.cf[1][0][i], fy = a.cf[1][1][j];
  const int sxi = slab_index(i, a.slab_a[0], a.slab_b[0]);
  const int syi = slab_index(j, a.slab_a[1], a.slab_b[1]);
  const bool extra = a.ndft || a.nprobe;
  for (int k = k0; k < k1; k++) {
    const long c = (long)i + sy * (long)j + sz * (long)k;
    const T ex = Ex[c], ey = Ey[c], ez = Ez[c];
    const T ez_yp = Ez[c + sy], ez_xp = Ez[c + 1];
    const T ey_xp = Ey[c + 1], ey_zp = Ey[c + sz];
    const T ex_yp = Ex[c + sy], ex_zp = Ex[c + sz];
    const T hx = Hx[c], hy = Hy[c], hz = Hz[c];
    const T im = mim[mat[c]];
    const T fz = a.cf[1][2][k];
    const T dxy = ez_yp - ez;  // Hx, J=y, +1
    const T dxz = ey_zp - ey;  // Hx, J=z, -1
    const T dyz = ex_zp - ex;  // Hy, J=z, +1
    const T dyx = ez_xp - ez;  // Hy, J=x, -1
    const T dzx = ey_xp - ey;  // Hz, J=x, +1
    const T dzy = ex_yp - ex;  // Hz, J=y, -1
    T d[3];
    d[0] = fz * dxz - fy * dxy;
    d[1] = fx * dyx - fz * dyz;
    d[2] = fy * dzy - fx * dzx;
    if (a.nsrc) add_sources(a, 2, c, d);
    if (sxi >= 0) {
      const long o = slab_off(a, 0, sxi, i, j, k);
      T* __restrict__ q1 = a.psi[1][1][0];
      T* __restrict__ q2 = a.psi[1][2][0];
      const T g = a.cg[1][0][i], b = a.cb[1][0][i], cc = a.cc[1][0][i];
      const T p1 = q1[o], p2 = q2[o];
      d[1] -= g * p1;
      d[2] -= g * p2;
      q1[o] = b * p1 - cc * dyx;
      q2[o] = b * p2 + cc * dzx;
    }
    if (syi >= 0) {
      const long o = slab_off(a, 1, syi, i, j, k);
      T* __restrict__ q0 = a.psi[1][0][1];
      T* __restrict__ q2 = a.psi[1][2][1];
      const T g = a.cg[1][1][j], b = a.cb[1][1][j], cc = a.cc[1][1][j];
      const T p0 = q0[o], p2 = q2[o];
      d[0] -= g * p0;
      d[2] -= g * p2;
      q0[o] = b * p0 + cc * dxy;
      q2[o] = b * p2 - cc * dzy;
    }
    const int szi = slab_index(k, a.slab_a[2], a.slab_b[2]);
    if (szi >= 0) {
      const long o = slab_off(a, 2, szi, i, j, k);
      T* __restrict__ q0 = a.psi[1][0][2];
      T* __restrict__ q1 = a.psi[1][1][2];
      const T g = a.cg[1][2][k], b = a.cb[1][2][k], cc = a.cc[1][2][k];
      const T p0 = q0[o], p1 = q1[o];
      d[0] -= g * p0;
      d[1] -= g * p1;
      q0[o] = b * p0 - cc * dxz;
      q1[o] = b * p1 + cc * dyz;
    }
    if (a.tfsf && (i == a.box[0] || i == a.box[1] || j == a.box[2] || j == a.box[3] || k == a.box[4] || k == a.box[5])) {
      const bool act[3] = {true, true, true};
      tfsf_cell<T, 3>(a, 1, i, j, k, act, act, d);
    }
    T h[3];
    h[0] = hx + d[0] * im;
    h[1] = hy + d[1] * im;
    h[2] = hz + d[2] * im;
    Hx[c] = h[0];
    Hy[c] = h[1];
    Hz[c] = h[2];
    if (extra) {
      const T e[3] = {ex, ey, ez};
      dft_probe<T>(a, c, i, j, k, e, h);
    }
  }
}

// ---------------------------------------------------------------- 2-D TM
// Ez, Hx, Hy (include/YeeFields.hpp:123-148); CPML along x and y.
template <typename T>
__global__ __launch_bounds__(256) void e_pass_tm(StepArgs<T> a, int) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= a.nx - 1 || j >= a.ny - 1) return;
  const T* __restrict__ Hx = a.H[0];
  const T* __restrict__ Hy = a.H[1];
  T* __restrict__ Ez = a.E[2];
  const long c = (long)i + a.sy * j;
  const T hx = Hx[c], hy = Hy[c], hy_xm = Hy[c - 1], hx_ym = Hx[c - a.sy];
  const T ez = Ez[c];
  const T ie = a.mie[a.mat[c]];
  const T dzx = hy - hy_xm;  // J=x, +1
  const T dzy = hx - hx_ym;  // J=y, -1
  T d[3] = {T(0), T(0), T(0)};
  d[2] = a.cf[0][0][i] * dzx - a.cf[0][1][j] * dzy;
  if (a.nsrc) add_sources(a, 0, c, d);
  const int sxi = slab_index(i, a.slab_a[0], a.slab_b[0]);
  if (sxi >= 0) {
    const long o = slab_off(a, 0, sxi, i, j, 0);
    T* __restrict__ q = a.psi[0][2][0];
    const T p = q[o];
    d[2] += a.cg[0][0][i] * p;
    q[o] = a.cb[0][0][i] * p + a.cc[0][0][i] * dzx;
  }
  const int syi = slab_index(j, a.slab_a[1], a.slab_b[1]);
  if (syi >= 0) {
    const long o = slab_off(a, 1, syi, i, j, 0);
    T* __restrict__ q = a.psi[0][2][1];
    const T p = q[o];
    d[2] += a.cg[0][1][j] * p;
    q[o] = a.cb[0][1][j] * p - a.cc[0][1][j] * dzy;
  }
  if (a.tfsf && (i == a.box[0] || i == a.box[1] || j == a.box[2] || j == a.box[3])) {
    const bool aF[3] = {false, false, true}, aC[3] = {true, true, false};
    tfsf_cell<T, 2>(a, 0, i, j, 0, aF, aC, d);
  }
  if (ie > T(0)) {
    Ez[c] = ez + d[2] * ie;
  } else if (ie < T(0) && a.Dd[2]) {
    const MatDev<T>& m = a.mats[a.mat[c]];
    Ez[c] = ade_component<T, 2>(a, m, dbox_index(a, i, j, 0), d[2]);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void h_pass_tm(StepArgs<T> a, int) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= a.nx - 1 || j >= a.ny - 1) return;
  const T* __restrict__ Ez = a.E[2];
  T* __restrict__ Hx = a.H[0];
  T* __restrict__ Hy = a.H[1];
  const long c = (long)i + a.sy * j;
  const T ez = Ez[c], ez_yp = Ez[c + a.sy], ez_xp = Ez[c + 1];
  const T hx = Hx[c], hy = Hy[c];
  const T im = a.mim[a.mat[c]];
  const T dxy = ez_yp - ez;  // Hx, J=y, +1
  const T dyx = ez_xp - ez;  // Hy, J=x, -1
  T d[3] = {T(0), T(0), T(0)};
  d[0] = -a.cf[1][1][j] * dxy;
  d[1] = a.cf[1][0][i] * dyx;
  if (a.nsrc) add_sources(a, 2, c, d);
  const int sxi = slab_index(i, a.slab_a[0], a.slab_b[0]);
  if (sxi >= 0) {
    const long o = slab_off(a, 0, sxi, i, j, 0);
    T* __restrict__ q = a.psi[1][1][0];
    const T p = q[o];
    d[1] -= a.cg[1][0][i] * p;
    q[o] = a.cb[1][0][i] * p - a.cc[1][0][i] * dyx;
  }
  const int syi = slab_index(j, a.slab_a[1], a.slab_b[1]);
  if (syi >= 0) {
    const long o = slab_off(a, 1, syi, i, j, 0);
    T* __restrict__ q = a.psi[1][0][1];
    const T p = q[o];
    d[0] -= a.cg[1][1][j] * p;
    q[o] = a.cb[1][1][j] * p + a.cc[1][1][j] * dxy;
  }
  if (a.tfsf && (i == a.box[0] || i == a.box[1] || j == a.box[2] || j == a.box[3])) {
    const bool aF[3] = {true, true, false}, aC[3] = {false, false, true};
    tfsf_cell<T, 2>(a, 1, i, j, 0, aF, aC, d);
  }
  T h[3] = {hx + d[0] * im, hy + d[1] * im, T(0)};
  Hx[c] = h[0];
  Hy[c] = h[1];
  if (a.ndft || a.nprobe) {
    const T e[3] = {T(0), T(0), ez};
    dft_probe<T>(a, c, i, j, 0, e, h);
  }
}

// ---------------------------------------------------------------- Aux1D line
// One workgroup advances the 1-D TEM incident line `nsteps` times
// (Aux1D::time_step, include/Source.hpp:413-435) and writes, per step, the
// line's Hy before the advance and its Ez after it.  Always FP64.
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

__global__ __launch_bounds__(1024) void aux_line(AuxArgs a) {
  const int L = a.L, tid = threadIdx.x, nt = blockDim.x;
  const double dt = a.dt, dx = a.dx;
  const int src = a.npml + 5;
  for (int s = 0; s < a.nsteps; s++) {
    for (int n = tid; n < L; n += nt) a.outH[(long)s * L + n] = a.Hy[n];
    __syncthreads();
    for (int n = tid; n < L; n += nt) {
      double hm = n > 0 ? a.Hy[n - 1] : 0.0;
      double diff = a.Hy[n] - hm;
      double D = a.Dz[n] + dt * (a.pE[4 * n + 2] / dx * diff);
      if (n < a.npml + 2 || n >= L - (a.npml + 3)) {
        D += dt * a.psiE[n] * a.pE[4 * n + 3];
        a.psiE[n] = a.pE[4 * n + 0] * a.psiE[n] + a.pE[4 * n + 1] / dx * diff;
      }
      a.Dz[n] = D;
      double Ev = D / a.eps0;
      long gs = a.step0 + s;
      if (n == src) Ev += gs < a.nsignal ? a.signal[gs] : 0.0;
      a.Ez[n] = Ev;
    }
    __syncthreads();
    for (int n = tid; n < L; n += nt) {
      double ep = n + 1 < L ? a.Ez[n + 1] : 0.0;
      double diff = ep - a.Ez[n];
      double Bv = a.By[n] + dt * (a.pH[4 * n + 2] / dx * diff);
      if (n < a.npml + 2 || n >= L - (a.npml + 3)) {
        Bv -= dt * a.psiH[n] * a.pH[4 * n + 3];
        a.psiH[n] = a.pH[4 * n + 0] * a.psiH[n] - a.pH[4 * n + 1] / dx * diff;
      }
      a.By[n] = Bv;
      a.outE[(long)s * L + n] = a.Ez[n];
    }
    __syncthreads();
    for (int n = tid; n < L; n += nt) a.Hy[n] = a.By[n] / a.mu0;
    __syncthreads();
  }
}

}  // namespace fdtdk
