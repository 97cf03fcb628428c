// Discrete adjoint (reverse-time) passes of the 3-D two-pass step, for the
// gradient of a running-DFT objective with respect to per-cell eps_r.
//
// The forward step n maps (E, H, psiE, psiH) to the next state through
//   psiE' = B psiE + C/dx sigma diff^-(H)         E' = E + ie (dt F/dx curl^-(H) + dt G psiE + s_n)
//   psiH' = B psiH + C/dx sigma diff^+(E')        H' = H - im (dt F/dx curl^+(E') + dt G psiH) + ...
// (e_cell_3d / h_cell_3d); sources and TFSF terms do not depend on the state.
// Its transpose, applied from the last step down to the first, is again a
// Yee-like pair of gather passes with the difference directions swapped:
//   adjoint E pass: nu += ie (objective source + curl^-T of w), w = dt F/dx mu - C/dx psiH-adjoint
//   adjoint H pass: mu += im curl^+T of v,                      v = dt F/dx nu + C/dx psiE-adjoint
// with the stored adjoint variables scaled as nu = ie * adj(E), mu = im * adj(H)
// so that no pass needs a neighbour's material.  The psiH adjoint is updated
// (own cell) in the adjoint H pass, the psiE adjoint (own cell) in the next
// adjoint E pass, so that no pass writes what another thread of it reads.
// d J / d eps_r = -(1/eps_r) sum_n adj(E^{n+1}) . (E^{n+1} - E^n) = -eps0 sum_n nu . dE.
// Cells of pole materials (Drude, Lorentz, conductive) carry D, P_p, J_p
// instead of E: their step D' = D + d, P_p' = P_p + dt J_p (conductive:
// (P_p + cD D')/(1 + cD)), E' = (D' - sum P_p') / (eps_r eps0),
// J_p' = cJ J_p + cE E' + cP P_p' is linear, and its transpose (adj_pole_cell)
// keeps the adjoints of D, P_p, J_p over the dispersive box and leaves in nu
// the adjoint of the curl increment d, which is what the adjoint H pass and
// the psiE adjoint read of every cell (ie adj(E) in a constant cell).
// The reference has no adjoint; this is the build's own.
#pragma once
#include "fdtd_kernels.hpp"

namespace fdtdk {

template <typename T>
struct AdjArgs {
  FDTD_G T* nu[3];          // ie * adjoint of E (pole cells: adjoint of the curl increment)
  FDTD_G T* mu[3];          // im * adjoint of H
  // adjoints of the pole cells' D, P_p, J_p over the dispersive box (null: none)
  FDTD_G T* aD[3];
  FDTD_G T* aP[kMaxPoles][3];
  FDTD_G T* aJ[kMaxPoles][3];
  FDTD_G T* lpsi[2][3][3];  // adjoint of psi[ft][I][J] (the forward slab layout)
  FDTD_G double* grad;      // [ndes]: d J / d eps_r of the design cells
  double eps0;
  // objective J = sum over the monitor's cells, slots and frequencies of |F|^2
  int obj_r[6];
  int obj_comp[3];  // E component of objective slot s (-1: unused)
  int obj_nfreq;
  long obj_ncells;
  const FDTD_G T* obj_w;        // [slot][nfreq][2][ncells]: 2 Re F, 2 Im F
  const FDTD_G double* obj_ph;  // E phases of the monitor, [nphase][nfreq][2]
  int obj_nphase;
  const FDTD_G AdjArgs* self;  // this block's own device address (null: none)
};

// The adjoint block as a cold branch reads it (the same device-address view as
// cold_args: loads in the branch stay in the branch).
template <typename T>
__device__ __forceinline__ const AdjArgs<T>& cold_adj(const AdjArgs<T>& q) {
  const FDTD_G AdjArgs<T>* p = q.self ? q.self : (const FDTD_G AdjArgs<T>*)&q;
  return (const AdjArgs<T>&)*p;
}

// objective source of E at (i,j,k) for the step's E sample
template <typename T>
__device__ FDTD_COLD void adj_objective(const AdjArgs<T>& q, long step, int i, int j, int k, T s[3]) {
  const int di = i - q.obj_r[0], dj = j - q.obj_r[2], dk = k - q.obj_r[4];
  const int wx = q.obj_r[1] - q.obj_r[0], wy = q.obj_r[3] - q.obj_r[2], wz = q.obj_r[5] - q.obj_r[4];
  if ((unsigned)di >= (unsigned)wx || (unsigned)dj >= (unsigned)wy || (unsigned)dk >= (unsigned)wz) return;
  if (step >= q.obj_nphase) return;
  const long lq = (long)di + (long)wx * ((long)dj + (long)wy * dk);
  for (int slot = 0; slot < 3; slot++) {
    const int comp = q.obj_comp[slot];
    if (comp < 0) continue;
    T acc = T(0);
    for (int f = 0; f < q.obj_nfreq; f++) {
      const FDTD_G double* ph = q.obj_ph + 2L * (q.obj_nfreq * step + f);
      const FDTD_G T* w = q.obj_w + ((long)slot * q.obj_nfreq + f) * 2 * q.obj_ncells + lq;
      acc += w[0] * (T)ph[0] + w[q.obj_ncells] * (T)ph[1];
    }
    if (comp == 0)
      s[0] += acc;
    else if (comp == 1)
      s[1] += acc;
    else
      s[2] += acc;
  }
}

// Reverse of one pole cell's material step, component by component: e is the
// adjoint of its new E gathered this step (objective and H pass); the D, P_p,
// J_p adjoints are read and written over the dispersive box; out receives the
// adjoint of the curl increment (= that of the new D).
template <typename T>
__device__ FDTD_COLD void adj_pole_cell(const StepArgs<T>& a0, const AdjArgs<T>& q0, int mid, long bq, const T e[3],
                                        T out[3]) {
  const StepArgs<T>& a = (const StepArgs<T>&)cold_args(a0);
  const AdjArgs<T>& q = cold_adj(q0);
  const MatDev<T>& m = a.mats[mid];
#pragma unroll
  for (int I = 0; I < 3; I++) {
    T eb = e[I];
    T jn[kMaxPoles], pt[kMaxPoles];
#pragma unroll
    for (int p = 0; p < kMaxPoles; p++) {
      if (p >= m.np) continue;
      const T jb = q.aJ[p][I][bq];
      eb += m.cE[p] * jb;  // J' = cJ J + cE E' + cP P'
      jn[p] = sizeof(T) == 4 ? jb - m.cJm[p] * jb : m.cJ[p] * jb;
      pt[p] = q.aP[p][I][bq] + m.cP[p] * jb;
    }
    const T nb = eb * m.inv_eps;  // E' = (D' - sum P') / (eps_r eps0)
    T db = q.aD[I][bq] + nb;
#pragma unroll
    for (int p = 0; p < kMaxPoles; p++) {
      if (p >= m.np) continue;
      pt[p] -= nb;
      if (m.cD[p] != T(0)) {  // P' = (P + cD D') / (1 + cD)
        const T t = pt[p] / (T(1) + m.cD[p]);
        db += m.cD[p] * t;
        pt[p] = t;
      } else {  // P' = P + dt J
        jn[p] += a.dt * pt[p];
      }
      q.aP[p][I][bq] = pt[p];
      q.aJ[p][I][bq] = jn[p];
    }
    q.aD[I][bq] = db;  // D' = D + d
    out[I] = db;
  }
}

// Adjoint E pass of step `step` over one column of cells.  hist: the step's
// row of design-cell E increments (null: no design box).  ADE: the grid has
// pole materials (their cells run adj_pole_cell).
// The column's slab indices and CPML coefficients along x and y are read
// once, before the march; the z coefficients come from zf, the block's rows
// of dt F/dx (E pass: H's of planes k0 - 1 .. k1 - 1; H pass: E's of planes
// k0 .. k1), so no uniform global load inside the march waits behind the
// field loads.
template <typename T, bool ADE = false>
__device__ __forceinline__ void adj_e_cells(const StepArgs<T>& a, const AdjArgs<T>& q, const T* tie, const T* zf,
                                            int i, int j, int k0, int k1, long step, const T* __restrict__ hist) {
  if (i < 1 || i >= a.nx - 1 || j >= a.ny - 1) return;
  const FDTD_G T* __restrict__ M0 = q.mu[0];
  const FDTD_G T* __restrict__ M1 = q.mu[1];
  const FDTD_G T* __restrict__ M2 = q.mu[2];
  FDTD_G T* __restrict__ N0 = q.nu[0];
  FDTD_G T* __restrict__ N1 = q.nu[1];
  FDTD_G T* __restrict__ N2 = q.nu[2];
  const long sy = a.sy, sz = a.sz;
  const T fx = a.cf[1][0][i], fy = a.cf[1][1][j];
  const T fxm = a.cf[1][0][i - 1], fym = a.cf[1][1][j - 1];
  const int sx0 = slab_index(i, a.slab_a[0], a.slab_b[0]), sx1 = slab_index(i - 1, a.slab_a[0], a.slab_b[0]);
  const int sy0 = slab_index(j, a.slab_a[1], a.slab_b[1]), sy1 = slab_index(j - 1, a.slab_a[1], a.slab_b[1]);
  T ccx = T(0), bx = T(0), gx = T(0), ccxm = T(0), ccy = T(0), by = T(0), gy = T(0), ccym = T(0);
  if (sx0 >= 0) {
    ccx = a.cc[1][0][i];
    bx = a.cb[0][0][i];
    gx = a.cg[0][0][i];
  }
  if (sx1 >= 0) ccxm = a.cc[1][0][i - 1];
  if (sy0 >= 0) {
    ccy = a.cc[1][1][j];
    by = a.cb[0][1][j];
    gy = a.cg[0][1][j];
  }
  if (sy1 >= 0) ccym = a.cc[1][1][j - 1];
  // whether the column meets the objective monitor and the design box (the
  // cold branches below are then taken only on their planes)
  const bool ocol = q.obj_w && i >= q.obj_r[0] && i < q.obj_r[1] && j >= q.obj_r[2] && j < q.obj_r[3];
  const bool dcol = hist && i >= a.dsg[0] && i < a.dsg[1] && j >= a.dsg[2] && j < a.dsg[3];
  // w12, w02 of the plane below (with their z-slab terms): the march carries
  // them from plane to plane instead of reloading that plane's mu
  T w12c, w02c;
  {
    const long c = (long)i + sy * (long)j + sz * (long)(k0 - 1);
    w12c = zf[0] * M1[c];
    w02c = zf[0] * M0[c];
    const int s = slab_index(k0 - 1, a.slab_a[2], a.slab_b[2]);
    if (s >= 0) {
      const long o = slab_off(a, 2, s, i, j, k0 - 1);
      const T cc = a.cc[1][2][k0 - 1];
      w02c -= cc * q.lpsi[1][0][2][o];
      w12c -= cc * q.lpsi[1][1][2][o];
    }
  }
  for (int k = k0; k < k1; k++) {
    const long c = (long)i + sy * (long)j + sz * (long)k;
    const T m0 = M0[c], m1 = M1[c], m2 = M2[c];
    const T m2_ym = M2[c - sy];
    const T m2_xm = M2[c - 1], m0_ym = M0[c - sy], m1_xm = M1[c - 1];
    T n0 = N0[c], n1 = N1[c], n2 = N2[c];
    const int mid = a.mat[c];
    const T ie = cell_inv_eps(a, tie, mid, i, j, k);
    const T fz = zf[k - k0 + 1];
    // w_KJ at the cell and at its lower J neighbour
    T w12 = fz * m1, w12m = w12c;
    T w21 = fy * m2, w21m = fym * m2_ym;
    T w02 = fz * m0, w02m = w02c;
    T w20 = fx * m2, w20m = fxm * m2_xm;
    T w01 = fy * m0, w01m = fym * m0_ym;
    T w10 = fx * m1, w10m = fxm * m1_xm;
    // x slabs: H psi (K=1,2; J=0) and the E psi adjoint (I=1,2; J=0) of the later step
    if (sx0 >= 0) {
      const long o = slab_off(a, 0, sx0, i, j, k);
      w10 -= ccx * q.lpsi[1][1][0][o];
      w20 -= ccx * q.lpsi[1][2][0][o];
      q.lpsi[0][1][0][o] = bx * q.lpsi[0][1][0][o] + gx * n1;
      q.lpsi[0][2][0][o] = bx * q.lpsi[0][2][0][o] + gx * n2;
    }
    if (sx1 >= 0) {
      const long o = slab_off(a, 0, sx1, i - 1, j, k);
      w10m -= ccxm * q.lpsi[1][1][0][o];
      w20m -= ccxm * q.lpsi[1][2][0][o];
    }
    if (sy0 >= 0) {
      const long o = slab_off(a, 1, sy0, i, j, k);
      w01 -= ccy * q.lpsi[1][0][1][o];
      w21 -= ccy * q.lpsi[1][2][1][o];
      q.lpsi[0][0][1][o] = by * q.lpsi[0][0][1][o] + gy * n0;
      q.lpsi[0][2][1][o] = by * q.lpsi[0][2][1][o] + gy * n2;
    }
    if (sy1 >= 0) {
      const long o = slab_off(a, 1, sy1, i, j - 1, k);
      w01m -= ccym * q.lpsi[1][0][1][o];
      w21m -= ccym * q.lpsi[1][2][1][o];
    }
    const int s = slab_index(k, a.slab_a[2], a.slab_b[2]);
    if (s >= 0) {
      const long o = slab_off(a, 2, s, i, j, k);
      const T cc = a.cc[1][2][k];
      w02 -= cc * q.lpsi[1][0][2][o];
      w12 -= cc * q.lpsi[1][1][2][o];
      const T b = a.cb[0][2][k], g = a.cg[0][2][k];
      q.lpsi[0][0][2][o] = b * q.lpsi[0][0][2][o] + g * n0;
      q.lpsi[0][1][2][o] = b * q.lpsi[0][1][2][o] + g * n1;
    }
    w12c = w12;
    w02c = w02;
    T src[3] = {T(0), T(0), T(0)};
    if (ocol && k >= q.obj_r[4] && k < q.obj_r[5]) adj_objective(cold_adj(q), step, i, j, k, src);
    if (ADE && ie < T(0)) {
      const T e[3] = {src[0] + (w12 - w12m) - (w21 - w21m), src[1] + (w20 - w20m) - (w02 - w02m),
                      src[2] + (w01 - w01m) - (w10 - w10m)};
      T o[3];
      adj_pole_cell(a, q, mid, dbox_index(a, i, j, k), e, o);
      n0 = o[0];
      n1 = o[1];
      n2 = o[2];
    } else {
      n0 += ie * (src[0] + (w12 - w12m) - (w21 - w21m));
      n1 += ie * (src[1] + (w20 - w20m) - (w02 - w02m));
      n2 += ie * (src[2] + (w01 - w01m) - (w10 - w10m));
    }
    N0[c] = n0;
    N1[c] = n1;
    N2[c] = n2;
    if (dcol && k >= a.dsg[4] && k < a.dsg[5]) {
      const StepArgs<T>& ac = (const StepArgs<T>&)cold_args(a);
      const AdjArgs<T>& qc = cold_adj(q);
      const long l = design_index(ac, i, j, k);
      if (l >= 0) {
        const long nd = (long)(ac.dsg[1] - ac.dsg[0]) * (ac.dsg[3] - ac.dsg[2]) * (ac.dsg[5] - ac.dsg[4]);
        const T dot = n0 * hist[l] + n1 * hist[nd + l] + n2 * hist[2 * nd + l];
        qc.grad[l] -= qc.eps0 * (double)dot;
      }
    }
  }
}

// Adjoint H pass of one step over one column of cells.
template <typename T>
__device__ __forceinline__ void adj_h_cells(const StepArgs<T>& a, const AdjArgs<T>& q, const T* tim, const T* zf, int i,
                                            int j, int k0, int k1) {
  if (i < 1 || i >= a.nx - 1 || j >= a.ny - 1) return;
  const FDTD_G T* __restrict__ N0 = q.nu[0];
  const FDTD_G T* __restrict__ N1 = q.nu[1];
  const FDTD_G T* __restrict__ N2 = q.nu[2];
  FDTD_G T* __restrict__ M0 = q.mu[0];
  FDTD_G T* __restrict__ M1 = q.mu[1];
  FDTD_G T* __restrict__ M2 = q.mu[2];
  const long sy = a.sy, sz = a.sz;
  const T fx = a.cf[0][0][i], fy = a.cf[0][1][j];
  const T fxp = a.cf[0][0][i + 1], fyp = a.cf[0][1][j + 1];
  const int sx0 = slab_index(i, a.slab_a[0], a.slab_b[0]), sx1 = slab_index(i + 1, a.slab_a[0], a.slab_b[0]);
  const int sy0 = slab_index(j, a.slab_a[1], a.slab_b[1]), sy1 = slab_index(j + 1, a.slab_a[1], a.slab_b[1]);
  T ccx = T(0), bx = T(0), gx = T(0), ccxp = T(0), ccy = T(0), by = T(0), gy = T(0), ccyp = T(0);
  if (sx0 >= 0) {
    ccx = a.cc[0][0][i];
    bx = a.cb[1][0][i];
    gx = a.cg[1][0][i];
  }
  if (sx1 >= 0) ccxp = a.cc[0][0][i + 1];
  if (sy0 >= 0) {
    ccy = a.cc[0][1][j];
    by = a.cb[1][1][j];
    gy = a.cg[1][1][j];
  }
  if (sy1 >= 0) ccyp = a.cc[0][1][j + 1];
  // the march runs down the planes and carries v12, v02 of the plane above
  // (with their z-slab terms) instead of reloading that plane's nu
  T v12c, v02c;
  {
    const long c = (long)i + sy * (long)j + sz * (long)k1;
    v12c = zf[k1 - k0] * N1[c];
    v02c = zf[k1 - k0] * N0[c];
    const int s = slab_index(k1, a.slab_a[2], a.slab_b[2]);
    if (s >= 0) {
      const long o = slab_off(a, 2, s, i, j, k1);
      const T cc = a.cc[0][2][k1];
      v02c += cc * q.lpsi[0][0][2][o];
      v12c += cc * q.lpsi[0][1][2][o];
    }
  }
  for (int k = k1 - 1; k >= k0; k--) {
    const long c = (long)i + sy * (long)j + sz * (long)k;
    const T n0 = N0[c], n1 = N1[c], n2 = N2[c];
    const T n2_yp = N2[c + sy];
    const T n2_xp = N2[c + 1], n0_yp = N0[c + sy], n1_xp = N1[c + 1];
    const T h0 = M0[c], h1 = M1[c], h2 = M2[c];
    const T im = tim[a.mat[c]];
    const T fz = zf[k - k0];
    // v_IJ at the cell and at its upper J neighbour
    T v12 = fz * n1, v12p = v12c;
    T v21 = fy * n2, v21p = fyp * n2_yp;
    T v02 = fz * n0, v02p = v02c;
    T v20 = fx * n2, v20p = fxp * n2_xp;
    T v01 = fy * n0, v01p = fyp * n0_yp;
    T v10 = fx * n1, v10p = fxp * n1_xp;
    // x slabs: E psi adjoint (I=1,2; J=0) and the H psi adjoint update (K=1,2; J=0)
    if (sx0 >= 0) {
      const long o = slab_off(a, 0, sx0, i, j, k);
      v10 += ccx * q.lpsi[0][1][0][o];
      v20 += ccx * q.lpsi[0][2][0][o];
      q.lpsi[1][1][0][o] = bx * q.lpsi[1][1][0][o] - gx * h1;
      q.lpsi[1][2][0][o] = bx * q.lpsi[1][2][0][o] - gx * h2;
    }
    if (sx1 >= 0) {
      const long o = slab_off(a, 0, sx1, i + 1, j, k);
      v10p += ccxp * q.lpsi[0][1][0][o];
      v20p += ccxp * q.lpsi[0][2][0][o];
    }
    if (sy0 >= 0) {
      const long o = slab_off(a, 1, sy0, i, j, k);
      v01 += ccy * q.lpsi[0][0][1][o];
      v21 += ccy * q.lpsi[0][2][1][o];
      q.lpsi[1][0][1][o] = by * q.lpsi[1][0][1][o] - gy * h0;
      q.lpsi[1][2][1][o] = by * q.lpsi[1][2][1][o] - gy * h2;
    }
    if (sy1 >= 0) {
      const long o = slab_off(a, 1, sy1, i, j + 1, k);
      v01p += ccyp * q.lpsi[0][0][1][o];
      v21p += ccyp * q.lpsi[0][2][1][o];
    }
    const int s = slab_index(k, a.slab_a[2], a.slab_b[2]);
    if (s >= 0) {
      const long o = slab_off(a, 2, s, i, j, k);
      const T cc = a.cc[0][2][k];
      v02 += cc * q.lpsi[0][0][2][o];
      v12 += cc * q.lpsi[0][1][2][o];
      const T b = a.cb[1][2][k], g = a.cg[1][2][k];
      q.lpsi[1][0][2][o] = b * q.lpsi[1][0][2][o] - g * h0;
      q.lpsi[1][1][2][o] = b * q.lpsi[1][1][2][o] - g * h1;
    }
    v12c = v12;
    v02c = v02;
    M0[c] = h0 + im * ((v12 - v12p) - (v21 - v21p));
    M1[c] = h1 + im * ((v20 - v20p) - (v02 - v02p));
    M2[c] = h2 + im * ((v01 - v01p) - (v10 - v10p));
  }
}

// launched like the forward passes: ntx x nty x nkc tiles (kc <= kAdjMaxKc),
// XCD-aware order, the H pass walking each XCD's range in reverse
constexpr int kAdjMaxKc = 64;
template <typename T, bool ADE>
__device__ __forceinline__ void adj_e_pass_body(const StepArgs<T>* __restrict__ ap, const AdjArgs<T>* __restrict__ qp,
                                                int kc, int ntx, int nty, int nkc, long step,
                                                const T* __restrict__ hist) {
  __shared__ MatLds<T> sm;
  __shared__ T zf[kAdjMaxKc + 1];
  const int t = xcd_tile(blockIdx.x, ntx * nty * nkc);
  const int bx = t % ntx, by = (t / ntx) % nty, bz = t / (ntx * nty);
  const int k0 = 1 + bz * kc, k1 = min(k0 + kc, ap->nz - 1);
  for (int x = threadIdx.x + blockDim.x * threadIdx.y; x <= k1 - k0; x += blockDim.x * blockDim.y)
    zf[x] = ap->cf[1][2][k0 - 1 + x];
  stage_materials(*ap, sm);  // (its barrier also covers zf)
  adj_e_cells<T, ADE>(*ap, *qp, sm.ie, (const T*)zf, bx * blockDim.x + threadIdx.x, 1 + by * blockDim.y + threadIdx.y,
                      k0, k1, step, hist);
}
template <typename T>
__device__ __forceinline__ void adj_h_pass_body(const StepArgs<T>* __restrict__ ap, const AdjArgs<T>* __restrict__ qp,
                                                int kc, int ntx, int nty, int nkc) {
  __shared__ MatLds<T> sm;
  __shared__ T zf[kAdjMaxKc + 1];
  const int t = xcd_tile_rev(blockIdx.x, ntx * nty * nkc);
  const int bx = t % ntx, by = (t / ntx) % nty, bz = t / (ntx * nty);
  const int k0 = 1 + bz * kc, k1 = min(k0 + kc, ap->nz - 1);
  for (int x = threadIdx.x + blockDim.x * threadIdx.y; x <= k1 - k0; x += blockDim.x * blockDim.y)
    zf[x] = ap->cf[0][2][k0 + x];
  stage_materials(*ap, sm);  // (its barrier also covers zf)
  adj_h_cells(*ap, *qp, sm.im, (const T*)zf, bx * blockDim.x + threadIdx.x, 1 + by * blockDim.y + threadIdx.y, k0,
              k1);
}

template <typename T, bool ADE = false>
__global__ __launch_bounds__(256) void adj_e_pass_3d(const StepArgs<T>* __restrict__ ap, const AdjArgs<T>* __restrict__ qp,
                                                     int kc, int ntx, int nty, int nkc, long step,
                                                     const T* __restrict__ hist) {
  adj_e_pass_body<T, ADE>(ap, qp, kc, ntx, nty, nkc, step, hist);
}
template <typename T>
__global__ __launch_bounds__(256) void adj_h_pass_3d(const StepArgs<T>* __restrict__ ap, const AdjArgs<T>* __restrict__ qp,
                                                     int kc, int ntx, int nty, int nkc) {
  adj_h_pass_body(ap, qp, kc, ntx, nty, nkc);
}

}  // namespace fdtdk
