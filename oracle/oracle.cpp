// oracle.cpp -- FP64 CPU restatement of LowMachFHD's implicit staggered Stokes solve.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the "port" CPU baseline.
// It follows the reference loop by loop on FBoxLib-style ghosted fabs of a single box that
// covers the domain.  Reference files (relative to staggered_grid/):
//   src_gmres/stag_applyop.f90        stag_applyop_2d/3d           -> applyop()
//   src_gmres/stag_mg_solver.f90      V-cycle, update, transfers   -> stag_mg_solver() etc.
//   src_gmres/stag_mg_layout.f90:179  compute_nlevs_mg             -> nlevs_mg()
//   src_gmres/apply_matrix.f90:24     apply_matrix                 -> apply_matrix()
//   src_gmres/apply_precon.f90:29     apply_precon (type 1)        -> apply_precon()
//   src_gmres/macproject.f90:34,318   macproject call site, subtract_weighted_gradp
//   src_gmres/gmres.f90:32            gmres, least_squares, rotmat, SolveUTriangular
//   src_gmres/norm_inner_product.f90  stag/cc inner products, sum_umac_press
//   src_common/div_and_grad.f90       compute_grad, compute_div
//   src_common/multifab_physbc_stag.f90 domainvel / macvel (homogeneous)
//   src_common/convert_stag.f90       average_cc_to_face / edge / node
//   src_lowMach/stochastic_m_fluxdiv.f90 stochastic_m_fluxdiv and helpers
// The cell-centred MAC multigrid (FBoxLib ml_cc_solve in the reference) is absent from the
// reference tree; mac_mg() is this build's own design and is mirrored exactly by the HIP code.
#include "oracle.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const int PERIODIC = -1;
inline bool is_noslip(int c) { return c >= 100 && c <= 103; }
inline bool is_slip(int c) { return c >= 200 && c <= 203; }
inline bool is_wall(int c) { return is_noslip(c) || is_slip(c); }

struct Geom {
  int dm;
  int n[3];
  double dx;
  int bclo[3], bchi[3];
  bool per(int d) const { return bclo[d] == PERIODIC; }
  long ncells() const { return (long)n[0] * n[1] * (dm == 3 ? n[2] : 1); }
  Geom coarsen(int f) const {
    Geom g = *this;
    for (int d = 0; d < dm; ++d) g.n[d] = n[d] / f;
    g.dx = dx * f;
    return g;
  }
};

// A fab covers [lo-ng, hi+ng] per direction; x fastest.
struct Fab {
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, ng[3] = {0, 0, 0};
  int sx = 1, sy = 1, sz = 1;
  std::vector<double> v;
  void define(const int* l, const int* h, const int* g) {
    for (int d = 0; d < 3; ++d) { lo[d] = l[d]; hi[d] = h[d]; ng[d] = g[d]; }
    sx = hi[0] - lo[0] + 1 + 2 * ng[0];
    sy = hi[1] - lo[1] + 1 + 2 * ng[1];
    sz = hi[2] - lo[2] + 1 + 2 * ng[2];
    v.assign((size_t)sx * sy * sz, 0.0);
  }
  inline double& operator()(int i, int j, int k) {
    return v[(size_t)(i - lo[0] + ng[0]) + (size_t)sx * ((j - lo[1] + ng[1]) + (size_t)sy * (k - lo[2] + ng[2]))];
  }
  inline double operator()(int i, int j, int k) const {
    return v[(size_t)(i - lo[0] + ng[0]) + (size_t)sx * ((j - lo[1] + ng[1]) + (size_t)sy * (k - lo[2] + ng[2]))];
  }
  inline double& at(const int* I) { return (*this)(I[0], I[1], I[2]); }
  inline double at(const int* I) const { return (*this)(I[0], I[1], I[2]); }
  void setval(double x) { std::fill(v.begin(), v.end(), x); }
  bool nodal(int d) const { return false; }
};

// nodal flags: which directions store n+1 points
Fab make_fab(const Geom& g, const bool* nodal, int ng) {
  int l[3] = {0, 0, 0}, h[3] = {0, 0, 0}, gg[3] = {0, 0, 0};
  for (int d = 0; d < g.dm; ++d) {
    h[d] = g.n[d] - 1 + (nodal[d] ? 1 : 0);
    gg[d] = ng;
  }
  Fab f;
  f.define(l, h, gg);
  return f;
}
Fab make_cc(const Geom& g, int ng) { bool nd[3] = {false, false, false}; return make_fab(g, nd, ng); }
Fab make_face(const Geom& g, int d, int ng) {
  bool nd[3] = {false, false, false};
  nd[d] = true;
  return make_fab(g, nd, ng);
}
// 3-D edges: 0 = xy (nodal x,y), 1 = xz, 2 = yz.  2-D: single nodal array (id 0).
Fab make_edge(const Geom& g, int e, int ng) {
  bool nd[3] = {false, false, false};
  if (g.dm == 2) { nd[0] = nd[1] = true; }
  else if (e == 0) { nd[0] = nd[1] = true; }
  else if (e == 1) { nd[0] = nd[2] = true; }
  else { nd[1] = nd[2] = true; }
  return make_fab(g, nd, ng);
}
inline int edge_id(int a, int b) {  // pair of directions -> edge array id (3-D)
  int lo = a < b ? a : b, hi = a < b ? b : a;
  if (lo == 0 && hi == 1) return 0;
  if (lo == 0 && hi == 2) return 1;
  return 2;
}

// ---------------------------------------------------------------------------------------
// valid-region <-> fab copies
void load_valid(Fab& f, const double* src) {
  size_t p = 0;
  for (int k = f.lo[2]; k <= f.hi[2]; ++k)
    for (int j = f.lo[1]; j <= f.hi[1]; ++j)
      for (int i = f.lo[0]; i <= f.hi[0]; ++i) f(i, j, k) = src[p++];
}
void store_valid(const Fab& f, double* dst) {
  size_t p = 0;
  for (int k = f.lo[2]; k <= f.hi[2]; ++k)
    for (int j = f.lo[1]; j <= f.hi[1]; ++j)
      for (int i = f.lo[0]; i <= f.hi[0]; ++i) dst[p++] = f(i, j, k);
}

// ---------------------------------------------------------------------------------------
// Ghost fills.
// multifab_fill_boundary for a single box: every ghost point whose periodic image lies in the
// valid region is overwritten by that image; valid points (including the duplicated nodal
// face at index n) are never touched.
void fill_boundary(Fab& f, const Geom& g) {
  int L[3], H[3];
  for (int d = 0; d < 3; ++d) { L[d] = f.lo[d] - f.ng[d]; H[d] = f.hi[d] + f.ng[d]; }
  for (int k = L[2]; k <= H[2]; ++k)
    for (int j = L[1]; j <= H[1]; ++j) {
      const bool row_interior = (j >= f.lo[1] && j <= f.hi[1]) && (k >= f.lo[2] && k <= f.hi[2]);
      for (int i = L[0]; i <= H[0]; ++i) {
        if (row_interior && i == f.lo[0]) i = f.hi[0] + 1;  // skip the valid part of the row
        if (i > H[0]) break;
        int I[3] = {i, j, k};
        bool ghost = false, ok = true;
        int S[3] = {i, j, k};
        for (int d = 0; d < g.dm; ++d) {
          if (I[d] < f.lo[d] || I[d] > f.hi[d]) {
            ghost = true;
            if (!g.per(d)) { ok = false; break; }
            S[d] = I[d] < f.lo[d] ? I[d] + g.n[d] : I[d] - g.n[d];
            if (S[d] < f.lo[d] || S[d] > f.hi[d]) { ok = false; break; }
          }
        }
        if (ghost && ok) f(i, j, k) = f(S[0], S[1], S[2]);
      }
    }
}

// copy the valid region of src into dst (fabs of the same index space, any ghost widths)
void copy_valid(Fab& dst, const Fab& src) {
  for (int k = src.lo[2]; k <= src.hi[2]; ++k)
    for (int j = src.lo[1]; j <= src.hi[1]; ++j)
      for (int i = src.lo[0]; i <= src.hi[0]; ++i) dst(i, j, k) = src(i, j, k);
}

// multifab_physbc_domainvel (homogeneous): normal velocity on wall faces and the ghost
// layers behind them is zero (multifab_physbc_stag.f90:271-387).
void domainvel(Fab& f, int d, const Geom& g) {
  for (int side = 0; side < 2; ++side) {
    int code = side == 0 ? g.bclo[d] : g.bchi[d];
    if (!is_wall(code)) continue;
    int a0 = side == 0 ? -f.ng[d] : g.n[d];
    int a1 = side == 0 ? 0 : g.n[d] + f.ng[d];
    int L[3], H[3];
    for (int e = 0; e < 3; ++e) { L[e] = f.lo[e]; H[e] = (e < g.dm) ? g.n[e] - 1 : f.hi[e]; }
    L[d] = a0; H[d] = a1;
    for (int k = L[2]; k <= H[2]; ++k)
      for (int j = L[1]; j <= H[1]; ++j)
        for (int i = L[0]; i <= H[0]; ++i) f(i, j, k) = 0.0;
  }
}

// multifab_physbc_macvel (homogeneous): transverse velocity ghost = -interior (no-slip,
// DIR_VEL) or +interior (slip, DIR_TRACT); walls processed x-lo, x-hi, y-lo, ... in order
// (multifab_physbc_stag.f90:1276-1645).
void macvel(Fab& f, int d, const Geom& g) {
  for (int e = 0; e < g.dm; ++e) {
    if (e == d) continue;
    for (int side = 0; side < 2; ++side) {
      int code = side == 0 ? g.bclo[e] : g.bchi[e];
      if (!is_wall(code)) continue;
      double sgn = is_noslip(code) ? -1.0 : 1.0;
      int L[3], H[3];
      for (int t = 0; t < 3; ++t) { L[t] = f.lo[t] - f.ng[t]; H[t] = f.hi[t] + f.ng[t]; }
      L[d] = 0; H[d] = g.n[d];          // the component's own nodal direction: valid only
      int src = side == 0 ? 0 : g.n[e] - 1;
      if (side == 0) { L[e] = -f.ng[e]; H[e] = -1; }
      else { L[e] = g.n[e]; H[e] = g.n[e] - 1 + f.ng[e]; }
      for (int k = L[2]; k <= H[2]; ++k)
        for (int j = L[1]; j <= H[1]; ++j)
          for (int i = L[0]; i <= H[0]; ++i) {
            int I[3] = {i, j, k};
            I[e] = src;
            f(i, j, k) = sgn * f.at(I);
          }
    }
  }
}

// full velocity ghost fill: domainvel, periodic fill, macvel (stag_mg_solver.f90:384-396)
void fill_vel(std::vector<Fab>& u, const Geom& g) {
  for (int d = 0; d < g.dm; ++d) {
    domainvel(u[d], d, g);
    fill_boundary(u[d], g);
    macvel(u[d], d, g);
  }
}

// Inhomogeneous velocity boundary values, as set_inhomogeneous_vel_bcs leaves them
// (multifab_physbc_stag.f90:2462-2549): n[a] a face fab of component a, read on its wall faces;
// t[q] the tangential values on the walls, a fab nodal in the wall direction d and the
// component b, q = d in 2-D and 2 d + (b > d ? b - 1 : b) in 3-D (the reference's order: y, z
// on x-walls, x, z on y-walls, x, y on z-walls). Slip-wall entries hold traction / eta.
struct VelBc {
  std::vector<Fab> n, t;
};
inline int vbt_index(int dm, int d, int b) { return dm == 2 ? d : 2 * d + (b > d ? b - 1 : b); }

// physbc_domainvel_{2d,3d}_inhomogeneous (:175-269, :389-533): the wall face and the ghost
// layers behind it take vel_bc_n at the wall face
void domainvel_inh(Fab& f, int d, const Geom& g, const Fab& vbn) {
  for (int side = 0; side < 2; ++side) {
    int code = side == 0 ? g.bclo[d] : g.bchi[d];
    if (!is_wall(code)) continue;
    int a0 = side == 0 ? -f.ng[d] : g.n[d];
    int a1 = side == 0 ? 0 : g.n[d] + f.ng[d];
    int L[3], H[3];
    for (int e = 0; e < 3; ++e) { L[e] = f.lo[e]; H[e] = (e < g.dm) ? g.n[e] - 1 : f.hi[e]; }
    L[d] = a0; H[d] = a1;
    for (int k = L[2]; k <= H[2]; ++k)
      for (int j = L[1]; j <= H[1]; ++j)
        for (int i = L[0]; i <= H[0]; ++i) {
          int W[3] = {i, j, k};
          W[d] = side == 0 ? 0 : g.n[d];
          f(i, j, k) = vbn.at(W);
        }
  }
}

// physbc_macvel_{2d,3d}_inhomogeneous (:1118-1274, :1647-2013): transverse ghosts, over the
// valid range of the other indices only, are 2 v - interior at a no-slip wall (DIR_VEL) and
// interior - dx v (lo side) / interior + dx v (hi side) at a slip wall (DIR_TRACT)
void macvel_inh(Fab& f, int d, const Geom& g, const std::vector<Fab>& vbt) {
  for (int e = 0; e < g.dm; ++e) {
    if (e == d) continue;
    const Fab& V = vbt[vbt_index(g.dm, e, d)];
    for (int side = 0; side < 2; ++side) {
      int code = side == 0 ? g.bclo[e] : g.bchi[e];
      if (!is_wall(code)) continue;
      const bool noslip = is_noslip(code);
      int L[3], H[3];
      for (int t = 0; t < 3; ++t) { L[t] = f.lo[t]; H[t] = f.hi[t]; }
      int src = side == 0 ? 0 : g.n[e] - 1;
      if (side == 0) { L[e] = -f.ng[e]; H[e] = -1; }
      else { L[e] = g.n[e]; H[e] = g.n[e] - 1 + f.ng[e]; }
      for (int k = L[2]; k <= H[2]; ++k)
        for (int j = L[1]; j <= H[1]; ++j)
          for (int i = L[0]; i <= H[0]; ++i) {
            int I[3] = {i, j, k}, W[3] = {i, j, k};
            I[e] = src;
            W[e] = side == 0 ? 0 : g.n[e];
            const double v = V.at(W);
            if (noslip) f(i, j, k) = 2.0 * v - f.at(I);
            else if (side == 0) f(i, j, k) = f.at(I) - g.dx * v;
            else f(i, j, k) = f.at(I) + g.dx * v;
          }
    }
  }
}

// apply_matrix's velocity fill when vel_bc_n and vel_bc_t are present (apply_matrix.f90:83-105)
void fill_vel_inh(std::vector<Fab>& u, const Geom& g, const VelBc& bc) {
  for (int d = 0; d < g.dm; ++d) {
    domainvel_inh(u[d], d, g, bc.n[d]);
    fill_boundary(u[d], g);
    macvel_inh(u[d], d, g, bc.t);
  }
}

// cell-centred ghost fill: periodic images, FOEXTRAP (copy of the adjacent interior cell)
// at walls (multifab_physbc.f90, pressure at walls is FOEXTRAP, define_bc_tower.f90:223)
void fill_cc(Fab& f, const Geom& g) {
  fill_boundary(f, g);
  for (int d = 0; d < g.dm; ++d)
    for (int side = 0; side < 2; ++side) {
      int code = side == 0 ? g.bclo[d] : g.bchi[d];
      if (!is_wall(code)) continue;
      int L[3], H[3];
      for (int t = 0; t < 3; ++t) { L[t] = f.lo[t] - f.ng[t]; H[t] = f.hi[t] + f.ng[t]; }
      int src = side == 0 ? 0 : g.n[d] - 1;
      if (side == 0) { L[d] = -f.ng[d]; H[d] = -1; }
      else { L[d] = g.n[d]; H[d] = g.n[d] - 1 + f.ng[d]; }
      for (int k = L[2]; k <= H[2]; ++k)
        for (int j = L[1]; j <= H[1]; ++j)
          for (int i = L[0]; i <= H[0]; ++i) {
            int I[3] = {i, j, k};
            I[d] = src;
            f(i, j, k) = f.at(I);
          }
    }
}

// ---------------------------------------------------------------------------------------
// Coefficients of one multigrid level.
struct Coef {
  Geom g;
  std::vector<Fab> alpha;          // dm face arrays, theta folded in
  Fab beta_own, gamma_own;         // storage for coarse levels (cc, 1 ghost)
  std::vector<Fab> bed_own;        // storage for coarse levels (edges / nodes, no ghosts)
  const Fab* beta = nullptr;       // cc, 1 ghost
  const Fab* gamma = nullptr;
  std::vector<const Fab*> bed;     // 3 edge arrays (3-D) or 1 nodal (2-D)
  void own() {
    beta = &beta_own; gamma = &gamma_own;
    bed.clear();
    for (auto& e : bed_own) bed.push_back(&e);
  }
};

// diagonal coefficient-sum and Lphi at face I of component a (stag_applyop.f90:517-1075,
// stag_applyop_2d 212-515; diagonal from stag_mg_update_2d/3d, stag_mg_solver.f90:1771-2458)
struct Stencil {
  int dm, visc;
  double dxsqinv;
  double b0, c0;  // constant-coefficient beta, gamma (visc_type > 0)
};

inline void off(int* I, int d, int s) { I[d] += s; }

double lphi_at(const Stencil& S, const Coef& C, const std::vector<Fab>& phi, int a, const int* I0,
               double* diag_out) {
  const int dm = S.dm;
  int T[2], nt = 0;
  for (int d = 0; d < dm; ++d) if (d != a) T[nt++] = d;
  int I[3] = {I0[0], I0[1], I0[2]};
  int Im[3] = {I0[0], I0[1], I0[2]};
  Im[a] -= 1;  // cell i-a
  const double pa = phi[a].at(I);
  int Ip[3] = {I0[0], I0[1], I0[2]}; Ip[a] += 1;
  const double pa_p = phi[a].at(Ip);
  const double pa_m = phi[a].at(Im);
  const double alpha = C.alpha[a].at(I);
  auto bedf = [&](int b) -> const Fab& { return *C.bed[dm == 2 ? 0 : edge_id(a, b)]; };
  double L, diag;
  if (S.visc < 0) {
    const double bI = C.beta->at(I), bM = C.beta->at(Im);
    const double gI = C.gamma->at(I), gM = C.gamma->at(Im);
    double nI, nM;  // normal-direction coefficient at cells i and i-a
    if (S.visc == -1) { nI = bI; nM = bM; }
    else if (S.visc == -2) { nI = 2.0 * bI; nM = 2.0 * bM; }
    else { nI = (4.0 / 3.0) * bI + gI; nM = (4.0 / 3.0) * bM + gM; }
    double dsum = nI + nM;
    double offs = pa_p * nI + pa_m * nM;
    double cross = 0.0;
    for (int t = 0; t < nt; ++t) {
      const int b = T[t];
      const Fab& E = bedf(b);
      int Ib[3] = {I0[0], I0[1], I0[2]}; Ib[b] += 1;
      int Ibm[3] = {I0[0], I0[1], I0[2]}; Ibm[b] -= 1;
      const double eI = E.at(I), eP = E.at(Ib);
      dsum += eI + eP;
      offs += phi[a].at(Ib) * eP + phi[a].at(Ibm) * eI;
      if (S.visc != -1) {
        int Imb[3] = {Im[0], Im[1], Im[2]}; Imb[b] += 1;  // i-a+b
        const Fab& pb = phi[b];
        if (S.visc == -2) {
          cross += pb.at(Ib) * eP - pb.at(I) * eI - pb.at(Imb) * eP + pb.at(Im) * eI;
        } else {
          const double tI = -(2.0 / 3.0) * bI + gI, tM = -(2.0 / 3.0) * bM + gM;
          cross += pb.at(Ib) * (eP + tI) - pb.at(I) * (eI + tI) - pb.at(Imb) * (eP + tM) +
                   pb.at(Im) * (eI + tM);
        }
      }
    }
    diag = alpha + dsum * S.dxsqinv;
    L = pa * diag - (offs + cross) * S.dxsqinv;
  } else {
    const double b = S.b0, c = S.c0;
    double nsum = 0.0, tsum = 0.0, cross = 0.0;
    nsum = pa_p + pa_m;
    for (int t = 0; t < nt; ++t) {
      const int bd = T[t];
      int Ib[3] = {I0[0], I0[1], I0[2]}; Ib[bd] += 1;
      int Ibm[3] = {I0[0], I0[1], I0[2]}; Ibm[bd] -= 1;
      tsum += phi[a].at(Ib) + phi[a].at(Ibm);
      if (S.visc != 1) {
        int Imb[3] = {Im[0], Im[1], Im[2]}; Imb[bd] += 1;
        const Fab& pb = phi[bd];
        cross += pb.at(Ib) - pb.at(I) - pb.at(Imb) + pb.at(Im);
      }
    }
    if (S.visc == 1) {
      diag = alpha + 2.0 * dm * b * S.dxsqinv;
      L = pa * diag - (nsum + tsum) * b * S.dxsqinv;
    } else if (S.visc == 2) {
      diag = alpha + (2.0 * dm + 2.0) * b * S.dxsqinv;
      L = pa * diag - (2.0 * nsum + tsum + cross) * b * S.dxsqinv;
    } else {
      const double num = dm == 3 ? 20.0 : 14.0;
      diag = alpha + (num * b * (1.0 / 3.0) + 2.0 * c) * S.dxsqinv;
      L = pa * diag -
          (nsum * ((4.0 / 3.0) * b + c) + tsum * b + cross * ((1.0 / 3.0) * b + c)) * S.dxsqinv;
    }
  }
  if (diag_out) *diag_out = diag;
  return L;
}

Stencil make_stencil(const Coef& C, int visc) {
  Stencil S;
  S.dm = C.g.dm;
  S.visc = visc;
  S.dxsqinv = 1.0 / (C.g.dx * C.g.dx);
  int z[3] = {0, 0, 0};
  S.b0 = C.beta->at(z);
  S.c0 = C.gamma->at(z);
  return S;
}

// diagonal of the operator at face I of component a (the "fac" of stag_mg_update)
double diag_at(const Stencil& S, const Coef& C, int a, const int* I0) {
  const int dm = S.dm;
  const int* I = I0;
  int Im[3] = {I0[0], I0[1], I0[2]};
  Im[a] -= 1;
  const double alpha = C.alpha[a].at(I);
  if (S.visc > 0) {
    if (S.visc == 1) return alpha + 2.0 * dm * S.b0 * S.dxsqinv;
    if (S.visc == 2) return alpha + (2.0 * dm + 2.0) * S.b0 * S.dxsqinv;
    const double num = dm == 3 ? 20.0 : 14.0;
    return alpha + (num / 3.0 * S.b0 + 2.0 * S.c0) * S.dxsqinv;
  }
  const double bI = C.beta->at(I), bM = C.beta->at(Im);
  double nI, nM;
  if (S.visc == -1) { nI = bI; nM = bM; }
  else if (S.visc == -2) { nI = 2.0 * bI; nM = 2.0 * bM; }
  else { nI = (4.0 / 3.0) * bI + C.gamma->at(I); nM = (4.0 / 3.0) * bM + C.gamma->at(Im); }
  double dsum = nI + nM;
  for (int b = 0; b < dm; ++b) {
    if (b == a) continue;
    const Fab& E = *C.bed[dm == 2 ? 0 : edge_id(a, b)];
    int Ib[3] = {I0[0], I0[1], I0[2]}; Ib[b] += 1;
    dsum += E.at(I) + E.at(Ib);
  }
  return alpha + dsum * S.dxsqinv;
}

inline bool color_has(int a, int color, int dm) {
  if (color == 0) return true;
  return (color - 1) / 2 == a;
}

// stag_applyop_level: Lphi on coloured faces, then zero on wall faces
void applyop(const Coef& C, int visc, const std::vector<Fab>& phi, std::vector<Fab>& Lphi, int color) {
  const Geom& g = C.g;
  Stencil S = make_stencil(C, visc);
  for (int a = 0; a < g.dm; ++a) {
    if (!color_has(a, color, g.dm)) continue;
    const int par = (color + 1) % 2;
    const int kh = g.dm == 3 ? g.n[2] - 1 + (a == 2) : 0;
    const int jh = g.n[1] - 1 + (a == 1);
    const int ih = g.n[0] - 1 + (a == 0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k <= kh; ++k)
      for (int j = 0; j <= jh; ++j) {
        int ioff = 0, step = 1;
        if (color != 0) { step = 2; if ((j + k) % 2 != par) ioff = 1; }
        for (int i = ioff; i <= ih; i += step) {
          int I[3] = {i, j, k};
          Lphi[a](i, j, k) = lphi_at(S, C, phi, a, I, nullptr);
        }
      }
  }
  // Lphi on physical boundaries set to zero, every component (stag_applyop.f90:203-206)
  for (int a = 0; a < g.dm; ++a) domainvel(Lphi[a], a, g);
}

// stag_mg_update: phi += omega (rhs - Lphi)/diag on coloured faces
void mg_update(const Coef& C, int visc, std::vector<Fab>& phi, const std::vector<Fab>& rhs,
               const std::vector<Fab>& Lphi, int color, double omega) {
  const Geom& g = C.g;
  Stencil S = make_stencil(C, visc);
  for (int a = 0; a < g.dm; ++a) {
    if (!color_has(a, color, g.dm)) continue;
    const int par = (color + 1) % 2;
    const int kh = g.dm == 3 ? g.n[2] - 1 + (a == 2) : 0;
    const int jh = g.n[1] - 1 + (a == 1);
    const int ih = g.n[0] - 1 + (a == 0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k <= kh; ++k)
      for (int j = 0; j <= jh; ++j) {
        int ioff = 0, step = 1;
        if (color != 0) { step = 2; if ((j + k) % 2 != par) ioff = 1; }
        for (int i = ioff; i <= ih; i += step) {
          int I[3] = {i, j, k};
          const double diag = diag_at(S, C, a, I);
          phi[a](i, j, k) = phi[a](i, j, k) + omega * (rhs[a](i, j, k) - Lphi[a](i, j, k)) / diag;
        }
      }
  }
}

// ---------------------------------------------------------------------------------------
// transfers (stag_mg_solver.f90:868-1628)
void cc_restriction(const Geom& gc, Fab& c, const Fab& f) {
  const double w = gc.dm == 3 ? 0.125 : 0.25;
  for (int k = 0; k < (gc.dm == 3 ? gc.n[2] : 1); ++k)
    for (int j = 0; j < gc.n[1]; ++j)
      for (int i = 0; i < gc.n[0]; ++i) {
        double s = 0.0;
        if (gc.dm == 3) {
          s = f(2 * i, 2 * j, 2 * k) + f(2 * i + 1, 2 * j, 2 * k) + f(2 * i, 2 * j + 1, 2 * k) +
              f(2 * i + 1, 2 * j + 1, 2 * k) + f(2 * i, 2 * j, 2 * k + 1) +
              f(2 * i + 1, 2 * j, 2 * k + 1) + f(2 * i, 2 * j + 1, 2 * k + 1) +
              f(2 * i + 1, 2 * j + 1, 2 * k + 1);
        } else {
          s = f(2 * i, 2 * j, 0) + f(2 * i + 1, 2 * j, 0) + f(2 * i, 2 * j + 1, 0) + f(2 * i + 1, 2 * j + 1, 0);
        }
        c(i, j, k) = w * s;
      }
  fill_boundary(c, gc);
}

// face restriction: simple (2/4-point mean, coefficients) or 6/12-point (residuals)
void stag_restriction(const Geom& gc, const std::vector<Fab>& f, std::vector<Fab>& c, bool simple) {
  const int dm = gc.dm;
  for (int a = 0; a < dm; ++a) {
    int T[2], nt = 0;
    for (int d = 0; d < dm; ++d) if (d != a) T[nt++] = d;
    const int kh = dm == 3 ? gc.n[2] - 1 + (a == 2) : 0;
    const int jh = gc.n[1] - 1 + (a == 1);
    const int ih = gc.n[0] - 1 + (a == 0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k <= kh; ++k)
      for (int j = 0; j <= jh; ++j)
        for (int i = 0; i <= ih; ++i) {
          int C0[3] = {2 * i, 2 * j, dm == 3 ? 2 * k : 0};
          // sum over the 2^(dm-1) fine faces on the coarse face, for normal offsets 0, +1, -1
          double s[3] = {0.0, 0.0, 0.0};
          for (int o = 0; o < 3; ++o) {
            const int no = o == 0 ? 0 : (o == 1 ? 1 : -1);
            if (simple && o > 0) break;
            for (int q = 0; q < (1 << nt); ++q) {
              int I[3] = {C0[0], C0[1], C0[2]};
              I[a] += no;
              for (int t = 0; t < nt; ++t) I[T[t]] += (q >> t) & 1;
              s[o] += f[a].at(I);
            }
          }
          double v;
          if (simple) v = (dm == 3 ? 0.25 : 0.5) * s[0];
          else if (dm == 3) v = 0.125 * s[0] + 0.0625 * s[1] + 0.0625 * s[2];
          else v = 0.25 * s[0] + 0.125 * (s[1] + s[2]);
          c[a](i, j, k) = v;
        }
  }
}

// edge (3-D) / nodal (2-D) coefficient coarsening
void edge_restriction(const Geom& gc, const std::vector<const Fab*>& fp, std::vector<Fab>& c) {
  const Fab* f[3] = {fp[0], fp.size() > 1 ? fp[1] : nullptr, fp.size() > 2 ? fp[2] : nullptr};
  if (gc.dm == 2) {
    for (int j = 0; j <= gc.n[1]; ++j)
      for (int i = 0; i <= gc.n[0]; ++i) c[0](i, j, 0) = (*f[0])(2 * i, 2 * j, 0);
    return;
  }
  // xy edges: mean along z; xz: along y; yz: along x
  for (int k = 0; k < gc.n[2]; ++k)
    for (int j = 0; j <= gc.n[1]; ++j)
      for (int i = 0; i <= gc.n[0]; ++i)
        c[0](i, j, k) = 0.5 * ((*f[0])(2 * i, 2 * j, 2 * k) + (*f[0])(2 * i, 2 * j, 2 * k + 1));
  for (int k = 0; k <= gc.n[2]; ++k)
    for (int j = 0; j < gc.n[1]; ++j)
      for (int i = 0; i <= gc.n[0]; ++i)
        c[1](i, j, k) = 0.5 * ((*f[1])(2 * i, 2 * j, 2 * k) + (*f[1])(2 * i, 2 * j + 1, 2 * k));
  for (int k = 0; k <= gc.n[2]; ++k)
    for (int j = 0; j <= gc.n[1]; ++j)
      for (int i = 0; i < gc.n[0]; ++i)
        c[2](i, j, k) = 0.5 * ((*f[2])(2 * i, 2 * j, 2 * k) + (*f[2])(2 * i + 1, 2 * j, 2 * k));
}

// prolongation-add, coarse ghosts filled (stag_mg_solver.f90:1365-1628)
void stag_prolongation(const Geom& gf, std::vector<Fab>& f, const std::vector<Fab>& c) {
  const int dm = gf.dm;
  for (int a = 0; a < dm; ++a) {
    int T[2], nt = 0;
    for (int d = 0; d < dm; ++d) if (d != a) T[nt++] = d;
    const int kh = dm == 3 ? gf.n[2] - 1 + (a == 2) : 0;
    const int jh = gf.n[1] - 1 + (a == 1);
    const int ih = gf.n[0] - 1 + (a == 0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k <= kh; ++k)
      for (int j = 0; j <= jh; ++j)
        for (int i = 0; i <= ih; ++i) {
          int I[3] = {i, j, k};
          int Cc[3] = {i / 2, j / 2, k / 2};
          int o[3];
          for (int d = 0; d < 3; ++d) o[d] = (I[d] % 2 == 0) ? -1 : 1;
          double add;
          if (dm == 3) {
            const int b = T[0], cdir = T[1];
            auto P = [&](int na, int nb, int nc) {
              int J[3] = {Cc[0], Cc[1], Cc[2]};
              J[a] += na; J[b] += nb; J[cdir] += nc;
              return c[a].at(J);
            };
            if (I[a] % 2 == 0) {
              add = (9.0 / 16.0) * P(0, 0, 0) + (3.0 / 16.0) * P(0, o[b], 0) +
                    (3.0 / 16.0) * P(0, 0, o[cdir]) + (1.0 / 16.0) * P(0, o[b], o[cdir]);
            } else {
              add = (9.0 / 32.0) * P(0, 0, 0) + (3.0 / 32.0) * P(0, o[b], 0) +
                    (3.0 / 32.0) * P(0, 0, o[cdir]) + (1.0 / 32.0) * P(0, o[b], o[cdir]) +
                    (9.0 / 32.0) * P(1, 0, 0) + (3.0 / 32.0) * P(1, o[b], 0) +
                    (3.0 / 32.0) * P(1, 0, o[cdir]) + (1.0 / 32.0) * P(1, o[b], o[cdir]);
            }
          } else {
            const int b = T[0];
            auto P = [&](int na, int nb) {
              int J[3] = {Cc[0], Cc[1], 0};
              J[a] += na; J[b] += nb;
              return c[a].at(J);
            };
            if (I[a] % 2 == 0) add = 0.75 * P(0, 0) + 0.25 * P(0, o[b]);
            else add = 0.375 * P(0, 0) + 0.125 * P(0, o[b]) + 0.375 * P(1, 0) + 0.125 * P(1, o[b]);
          }
          f[a](i, j, k) = f[a](i, j, k) + add;
        }
  }
}

// compute_nlevs_mg (stag_mg_layout.f90:179-222)
int nlevs_mg(const Geom& g, int minwidth) {
  int nl = -1;
  for (int d = 0; d < g.dm; ++d) {
    int t = g.n[d], r = 1;
    while (t % 2 == 0 && t / minwidth != 1) { t /= 2; ++r; }
    nl = nl < 0 ? r : std::min(nl, r);
  }
  return nl;
}

// ---------------------------------------------------------------------------------------
struct Ctx {
  orc_params p;
  Geom g;
  std::vector<Fab> alpha;  // unscaled alpha_fc (ng 0)
  Fab beta, gamma;         // ng 1
  std::vector<Fab> bed;
  double theta = 0.0;
  bool have_coef = false;
};

std::vector<Fab> make_vel(const Geom& g, int ng) {
  std::vector<Fab> u;
  for (int d = 0; d < g.dm; ++d) u.push_back(make_face(g, d, ng));
  return u;
}

double inf_norm(const Fab& f) {
  double m = 0.0;
  for (int k = f.lo[2]; k <= f.hi[2]; ++k)
    for (int j = f.lo[1]; j <= f.hi[1]; ++j)
      for (int i = f.lo[0]; i <= f.hi[0]; ++i) m = std::max(m, std::fabs(f(i, j, k)));
  return m;
}

// stag_mg_solver (stag_mg_solver.f90:37-864), single box: the fancy bottom solver is
// inactive (n_cells/max_grid_size = 1), so the bottom is nu_bottom smoothing sweeps.
void stag_mg_solver(Ctx& X, std::vector<Fab>& phi_fc, const std::vector<Fab>& rhs_fc) {
  const orc_params& p = X.p;
  const Geom& g0 = X.g;
  const int dm = g0.dm;
  const int nl = nlevs_mg(g0, p.stag_minwidth);
  std::vector<Coef> C(nl);
  std::vector<std::vector<Fab>> rhs(nl), phi(nl), Lphi(nl);
  for (int n = 0; n < nl; ++n) {
    Geom g = g0.coarsen(1 << n);
    C[n].g = g;
    C[n].alpha = make_vel(g, 0);
    if (n > 0) {
      C[n].beta_own = make_cc(g, 1);
      C[n].gamma_own = make_cc(g, 1);
      for (int e = 0; e < (dm == 3 ? 3 : 1); ++e) C[n].bed_own.push_back(make_edge(g, e, 0));
      C[n].own();
    }
    rhs[n] = make_vel(g, 0);
    phi[n] = make_vel(g, 1);
    Lphi[n] = make_vel(g, 1);
  }
  // level-1 coefficients, alpha *= theta (stag_mg_solver.f90:200-212)
  C[0].beta = &X.beta;
  C[0].gamma = &X.gamma;
  for (auto& e : X.bed) C[0].bed.push_back(&e);
  for (int d = 0; d < dm; ++d) {
    C[0].alpha[d].v = X.alpha[d].v;
    for (auto& x : C[0].alpha[d].v) x *= X.theta;
  }
  for (int n = 1; n < nl; ++n) {  // coarsen coefficients (:214-230)
    cc_restriction(C[n].g, C[n].beta_own, *C[n - 1].beta);
    cc_restriction(C[n].g, C[n].gamma_own, *C[n - 1].gamma);
    stag_restriction(C[n].g, C[n - 1].alpha, C[n].alpha, true);
    edge_restriction(C[n].g, C[n - 1].bed, C[n].bed_own);
  }
  for (int d = 0; d < dm; ++d) {
    copy_valid(phi[0][d], phi_fc[d]);
    copy_valid(rhs[0][d], rhs_fc[d]);
  }
  fill_vel(phi[0], C[0].g);

  // initial residual (:259-313); only its "all zero" test changes the result
  applyop(C[0], p.visc_type, phi[0], Lphi[0], 0);
  double resid0[3] = {0, 0, 0};
  for (int d = 0; d < dm; ++d) {
    double m = 0.0;
    const Fab& L = Lphi[0][d];
    const Fab& R = rhs[0][d];
    for (int k = R.lo[2]; k <= R.hi[2]; ++k)
      for (int j = R.lo[1]; j <= R.hi[1]; ++j)
        for (int i = R.lo[0]; i <= R.hi[0]; ++i) m = std::max(m, std::fabs(L(i, j, k) - R(i, j, k)));
    resid0[d] = m;
  }
  bool allzero = true;
  for (int d = 0; d < dm; ++d) if (resid0[d] != 0.0) allzero = false;
  if (allzero) return;  // phi_fc unchanged

  const int c0 = p.stag_smoother == 0 ? 0 : 1;
  const int c1 = p.stag_smoother == 0 ? 0 : 2 * dm;
  auto smooth = [&](int n, int nsweeps) {
    for (int m = 0; m < nsweeps; ++m)
      for (int color = c0; color <= c1; ++color) {
        applyop(C[n], p.visc_type, phi[n], Lphi[n], color);
        mg_update(C[n], p.visc_type, phi[n], rhs[n], Lphi[n], color, p.stag_omega);
        fill_vel(phi[n], C[n].g);
      }
  };
  auto residual = [&](int n) {  // Lphi <- rhs - L phi, with BC fill (:403-423)
    applyop(C[n], p.visc_type, phi[n], Lphi[n], 0);
    for (int d = 0; d < dm; ++d) {
      Fab& L = Lphi[n][d];
      const Fab& R = rhs[n][d];
      for (int k = R.lo[2]; k <= R.hi[2]; ++k)
        for (int j = R.lo[1]; j <= R.hi[1]; ++j)
          for (int i = R.lo[0]; i <= R.hi[0]; ++i) L(i, j, k) = -(L(i, j, k) - R(i, j, k));
    }
    fill_vel(Lphi[n], C[n].g);
  };

  for (int vc = 0; vc < p.stag_max_vcycles; ++vc) {
    for (int n = 1; n < nl; ++n)
      for (int d = 0; d < dm; ++d) phi[n][d].setval(0.0);
    for (int n = 0; n < nl - 1; ++n) {
      smooth(n, p.stag_nu_down);
      residual(n);
      stag_restriction(C[n + 1].g, Lphi[n], rhs[n + 1], false);
      for (int d = 0; d < dm; ++d) domainvel(rhs[n + 1][d], d, C[n + 1].g);
    }
    smooth(nl - 1, p.stag_nu_bottom);
    for (int n = nl - 2; n >= 0; --n) {
      stag_prolongation(C[n].g, phi[n], phi[n + 1]);
      fill_vel(phi[n], C[n].g);
      smooth(n, p.stag_nu_up);
    }
    // convergence test (:761-807)
    applyop(C[0], p.visc_type, phi[0], Lphi[0], 0);
    bool done = true;
    for (int d = 0; d < dm; ++d) {
      double m = 0.0;
      const Fab& L = Lphi[0][d];
      const Fab& R = rhs[0][d];
      for (int k = R.lo[2]; k <= R.hi[2]; ++k)
        for (int j = R.lo[1]; j <= R.hi[1]; ++j)
          for (int i = R.lo[0]; i <= R.hi[0]; ++i) m = std::max(m, std::fabs(L(i, j, k) - R(i, j, k)));
      double r0 = resid0[d];
      if (r0 == 0.0) { r0 = std::max(resid0[0], std::max(resid0[1], dm == 3 ? resid0[2] : 0.0)); }
      if (!(m / r0 < p.stag_rel_tol)) done = false;
    }
    if (done) break;
  }
  for (int d = 0; d < dm; ++d) copy_valid(phi_fc[d], phi[0][d]);
  fill_vel(phi_fc, g0);
}

// ---------------------------------------------------------------------------------------
// compute_grad (div_and_grad.f90:417-703): face gradient of a cc field with filled ghosts;
// x2 on wall faces because the ghost holds the boundary value (FOEXTRAP)
void compute_grad(const Geom& g, const Fab& p, std::vector<Fab>& gp) {
  const double dxinv = 1.0 / g.dx;
  for (int a = 0; a < g.dm; ++a) {
    Fab& G = gp[a];
    const int kh = g.dm == 3 ? g.n[2] - 1 + (a == 2) : 0;
    const int jh = g.n[1] - 1 + (a == 1);
    const int ih = g.n[0] - 1 + (a == 0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k <= kh; ++k)
      for (int j = 0; j <= jh; ++j)
        for (int i = 0; i <= ih; ++i) {
          int I[3] = {i, j, k}, M[3] = {i, j, k};
          M[a] -= 1;
          double fac = dxinv;
          if ((I[a] == 0 && is_wall(g.bclo[a])) || (I[a] == g.n[a] && is_wall(g.bchi[a]))) fac = 2.0 * dxinv;
          G(i, j, k) = (p.at(I) - p.at(M)) * fac;
        }
  }
}

// compute_div (div_and_grad.f90:884-1035)
void compute_div(const Geom& g, const std::vector<Fab>& u, Fab& div) {
  const double dxinv = 1.0 / g.dx;
  const int kh = g.dm == 3 ? g.n[2] - 1 : 0;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k <= kh; ++k)
    for (int j = 0; j < g.n[1]; ++j)
      for (int i = 0; i < g.n[0]; ++i) {
        double s = (u[0](i + 1, j, k) - u[0](i, j, k)) * dxinv + (u[1](i, j + 1, k) - u[1](i, j, k)) * dxinv;
        if (g.dm == 3) s += (u[2](i, j, k + 1) - u[2](i, j, k)) * dxinv;
        div(i, j, k) = s;
      }
}

// ---------------------------------------------------------------------------------------
// Cell-centred MAC multigrid (own design; FBoxLib's ml_cc_solve is not in the reference).
// Operator: (L phi)_c = sum_d [b(c+e_d)(phi(c+e_d)-phi(c)) - b(c)(phi(c)-phi(c-e_d))]/h^2
// with b = 1/alpha_fc on faces and b = 0 on wall faces (homogeneous Neumann).
// Red-black Gauss-Seidel (even parity first; each colour computed from the values before
// the colour), nu_down/nu_up sweeps, full-weighting (mean of 2^dm children) restriction,
// piecewise-constant prolongation, coarse face coefficient = mean of the 2^(dm-1) fine faces
// on it, bottom: mean of rhs removed (singular problem), nu_bottom sweeps.
struct MacLev {
  Geom g;
  std::vector<Fab> b;  // face coefficients (ng 0)
  Fab phi, rhs, res;   // phi ng 1
};

double mac_L(const MacLev& M, int i, int j, int k, double* diag) {
  const Geom& g = M.g;
  const double h2i = 1.0 / (g.dx * g.dx);
  const double pc = M.phi(i, j, k);
  double s = 0.0, dsum = 0.0;
  for (int d = 0; d < g.dm; ++d) {
    int I[3] = {i, j, k}, P[3] = {i, j, k}, Mi[3] = {i, j, k};
    P[d] += 1; Mi[d] -= 1;
    const double bp = M.b[d].at(P), bm = M.b[d].at(I);
    s += bp * (M.phi.at(P) - pc) - bm * (pc - M.phi.at(Mi));
    dsum += bp + bm;
  }
  if (diag) *diag = -dsum * h2i;
  return s * h2i;
}

void mac_gs(MacLev& M, int nsweeps) {
  const Geom& g = M.g;
  const int nz = g.dm == 3 ? g.n[2] : 1;
  std::vector<double> nv((size_t)g.n[0] * g.n[1] * nz);
  for (int s = 0; s < nsweeps; ++s)
    for (int par = 0; par < 2; ++par) {
      fill_cc(M.phi, g);
#pragma omp parallel for collapse(2) schedule(static)
      for (int k = 0; k < nz; ++k)
        for (int j = 0; j < g.n[1]; ++j)
          for (int i = 0; i < g.n[0]; ++i) {
            if (((i + j + k) & 1) != par) continue;
            double diag;
            const double L = mac_L(M, i, j, k, &diag);
            double v = M.phi(i, j, k);
            if (diag != 0.0) v += (M.rhs(i, j, k) - L) / diag;
            nv[(size_t)i + (size_t)g.n[0] * (j + (size_t)g.n[1] * k)] = v;
          }
      for (int k = 0; k < nz; ++k)
        for (int j = 0; j < g.n[1]; ++j)
          for (int i = 0; i < g.n[0]; ++i)
            if (((i + j + k) & 1) == par) M.phi(i, j, k) = nv[(size_t)i + (size_t)g.n[0] * (j + (size_t)g.n[1] * k)];
    }
  fill_cc(M.phi, g);
}

// the MAC solve with level-0 face coefficient b0 (walls zero) when given, else 1/alpha
void mac_mg(Ctx& X, const Fab& rhs_in, Fab& phi_out, const std::vector<Fab>* b0 = nullptr) {
  const orc_params& p = X.p;
  const Geom& g0 = X.g;
  const int dm = g0.dm;
  const int nl = nlevs_mg(g0, p.mac_minwidth);
  std::vector<MacLev> L(nl);
  for (int n = 0; n < nl; ++n) {
    L[n].g = g0.coarsen(1 << n);
    L[n].b = make_vel(L[n].g, 0);
    L[n].phi = make_cc(L[n].g, 1);
    L[n].rhs = make_cc(L[n].g, 0);
    L[n].res = make_cc(L[n].g, 0);
  }
  for (int d = 0; d < dm; ++d) {
    Fab& B = L[0].b[d];
    const Fab& A = X.alpha[d];
    if (b0) B.v = (*b0)[d].v;
    else
      for (size_t q = 0; q < B.v.size(); ++q) B.v[q] = 1.0 / A.v[q];
    // Neumann: zero coefficient on wall faces
    for (int side = 0; side < 2; ++side) {
      if (!is_wall(side == 0 ? g0.bclo[d] : g0.bchi[d])) continue;
      int idx = side == 0 ? 0 : g0.n[d];
      for (int k = B.lo[2]; k <= B.hi[2]; ++k)
        for (int j = B.lo[1]; j <= B.hi[1]; ++j)
          for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
            int I[3] = {i, j, k};
            if (I[d] == idx) B(i, j, k) = 0.0;
          }
    }
  }
  for (int n = 1; n < nl; ++n) stag_restriction(L[n].g, L[n - 1].b, L[n].b, true);
  for (int k = 0; k <= L[0].rhs.hi[2]; ++k)
    for (int j = 0; j <= L[0].rhs.hi[1]; ++j)
      for (int i = 0; i <= L[0].rhs.hi[0]; ++i) L[0].rhs(i, j, k) = rhs_in(i, j, k);

  bool singular = true;  // no Dirichlet condition on the pressure anywhere
  for (int vc = 0; vc < p.mac_max_vcycles; ++vc) {
    for (int n = 1; n < nl; ++n) L[n].phi.setval(0.0);
    for (int n = 0; n < nl - 1; ++n) {
      mac_gs(L[n], p.mac_nu_down);
      MacLev& M = L[n];
      const int nz = M.g.dm == 3 ? M.g.n[2] : 1;
      for (int k = 0; k < nz; ++k)
        for (int j = 0; j < M.g.n[1]; ++j)
          for (int i = 0; i < M.g.n[0]; ++i) M.res(i, j, k) = M.rhs(i, j, k) - mac_L(M, i, j, k, nullptr);
      MacLev& Cc = L[n + 1];
      const int nzc = Cc.g.dm == 3 ? Cc.g.n[2] : 1;
      const double w = dm == 3 ? 0.125 : 0.25;
      for (int k = 0; k < nzc; ++k)
        for (int j = 0; j < Cc.g.n[1]; ++j)
          for (int i = 0; i < Cc.g.n[0]; ++i) {
            double s = 0.0;
            for (int q = 0; q < (1 << dm); ++q)
              s += M.res(2 * i + (q & 1), 2 * j + ((q >> 1) & 1), dm == 3 ? 2 * k + ((q >> 2) & 1) : 0);
            Cc.rhs(i, j, k) = w * s;
          }
    }
    MacLev& B = L[nl - 1];
    if (singular) {
      double s = 0.0;
      for (double x : B.rhs.v) s += x;
      const double mean = s / (double)B.g.ncells();
      for (double& x : B.rhs.v) x -= mean;
    }
    mac_gs(B, p.mac_nu_bottom);
    for (int n = nl - 2; n >= 0; --n) {
      MacLev& F = L[n];
      const MacLev& Cc = L[n + 1];
      const int nz = F.g.dm == 3 ? F.g.n[2] : 1;
      for (int k = 0; k < nz; ++k)
        for (int j = 0; j < F.g.n[1]; ++j)
          for (int i = 0; i < F.g.n[0]; ++i) F.phi(i, j, k) += Cc.phi(i / 2, j / 2, k / 2);
      mac_gs(F, p.mac_nu_up);
    }
  }
  for (int k = 0; k <= L[0].rhs.hi[2]; ++k)
    for (int j = 0; j <= L[0].rhs.hi[1]; ++j)
      for (int i = 0; i <= L[0].rhs.hi[0]; ++i) phi_out(i, j, k) = L[0].phi(i, j, k);
  fill_cc(phi_out, g0);
}

static void mac_apply(Ctx& X, const Fab& phi_in, Fab& out) {
  const Geom& g = X.g;
  MacLev M;
  M.g = g;
  M.b = make_vel(g, 0);
  for (int d = 0; d < g.dm; ++d) {
    Fab& B = M.b[d];
    const Fab& A = X.alpha[d];
    for (size_t q = 0; q < B.v.size(); ++q) B.v[q] = 1.0 / A.v[q];
    for (int side = 0; side < 2; ++side) {
      if (!is_wall(side == 0 ? g.bclo[d] : g.bchi[d])) continue;
      const int ia = side == 0 ? 0 : g.n[d];
      for (int k = B.lo[2]; k <= B.hi[2]; ++k)
        for (int j = B.lo[1]; j <= B.hi[1]; ++j)
          for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
            const int I[3] = {i, j, k};
            if (I[d] == ia) B(i, j, k) = 0.0;
          }
    }
  }
  M.phi = make_cc(g, 1);
  copy_valid(M.phi, phi_in);
  fill_cc(M.phi, g);
  for (int k = 0; k <= out.hi[2]; ++k)
    for (int j = 0; j <= out.hi[1]; ++j)
      for (int i = 0; i <= out.hi[0]; ++i) out(i, j, k) = mac_L(M, i, j, k, nullptr);
}

// ---------------------------------------------------------------------------------------
// apply_matrix (apply_matrix.f90:24-157): x_u, x_p get their ghosts (and wall values) set
// stag_applyop_4o (stag_applyop.f90:1262-1510), visc_type 1 on a periodic box: the
// 4th-order Laplacian of each component, (theta alpha + K b/(12 h^2)) phi - (16 phi(i+1) +
// 16 phi(i-1) - phi(i+2) - phi(i-2) + the same in j and k) b/(12 h^2), K = 60 (2-D) / 90
// (3-D); neighbours beyond the valid region are periodic images (index + n below 0, - n
// above the last valid index: face n is its own degree of freedom)
void applyop_4o(const Geom& g, const std::vector<Fab>& alpha_theta, double b, const std::vector<Fab>& phi,
                std::vector<Fab>& L) {
  const double t12 = 1.0 / (12.0 * (g.dx * g.dx));
  const double K = g.dm == 3 ? 90.0 : 60.0;
  for (int a = 0; a < g.dm; ++a) {
    const Fab& P = phi[a];
    Fab& R = L[a];
    for (int k = R.lo[2]; k <= R.hi[2]; ++k)
      for (int j = R.lo[1]; j <= R.hi[1]; ++j)
        for (int i = R.lo[0]; i <= R.hi[0]; ++i) {
          auto v = [&](int di, int dj, int dk) {
            int I[3] = {i + di, j + dj, k + dk};
            for (int d = 0; d < 3; ++d) {
              const int top = g.n[d] - 1 + (d == a);
              if (I[d] < 0) I[d] += g.n[d];
              else if (I[d] > top) I[d] -= g.n[d];
            }
            return P(I[0], I[1], I[2]);
          };
          double s = 16.0 * v(1, 0, 0);
          s = s + 16.0 * v(-1, 0, 0);
          s = s - 1.0 * v(2, 0, 0);
          s = s - 1.0 * v(-2, 0, 0);
          s = s + 16.0 * v(0, 1, 0);
          s = s + 16.0 * v(0, -1, 0);
          s = s - 1.0 * v(0, 2, 0);
          s = s - 1.0 * v(0, -2, 0);
          if (g.dm == 3) {
            s = s + 16.0 * v(0, 0, 1);
            s = s + 16.0 * v(0, 0, -1);
            s = s - 1.0 * v(0, 0, 2);
            s = s - 1.0 * v(0, 0, -2);
          }
          R(i, j, k) = (alpha_theta[a](i, j, k) + K * b * t12) * P(i, j, k) - s * b * t12;
        }
  }
}

// compute_grad_4o (div_and_grad.f90:705-880) as written there: (-p(i+1) + 15 p(i) - 15 p(i-1)
// + p(i-2)) / (12 h) on every face, periodic cell images (a second-order formula despite the
// name: the fourth-order staggered weights would be (-1, 27, -27, 1)/24)
void compute_grad_4o(const Geom& g, const Fab& p, std::vector<Fab>& gp) {
  const double t = 1.0 / (12.0 * g.dx);
  for (int a = 0; a < g.dm; ++a) {
    Fab& R = gp[a];
    for (int k = R.lo[2]; k <= R.hi[2]; ++k)
      for (int j = R.lo[1]; j <= R.hi[1]; ++j)
        for (int i = R.lo[0]; i <= R.hi[0]; ++i) {
          auto c = [&](int s) {
            int I[3] = {i, j, k};
            I[a] += s;
            for (int d = 0; d < 3; ++d) {
              if (I[d] < 0) I[d] += g.n[d];
              else if (I[d] >= g.n[d]) I[d] -= g.n[d];
            }
            return p(I[0], I[1], I[2]);
          };
          R(i, j, k) = (-c(1) + 15.0 * c(0) - 15.0 * c(-1) + c(-2)) * t;
        }
  }
}

void apply_matrix(Ctx& X, std::vector<Fab>& x_u, Fab& x_p, std::vector<Fab>& b_u, Fab& b_p,
                  const VelBc* bc = nullptr) {
  const Geom& g = X.g;
  fill_cc(x_p, g);
  if (bc) fill_vel_inh(x_u, g, *bc);
  else fill_vel(x_u, g);
  if (X.p.gmres_spatial_order == 4) {  // apply_matrix.f90:123-135
    std::vector<Fab> at = make_vel(g, 0);
    for (int d = 0; d < g.dm; ++d) {
      at[d].v = X.alpha[d].v;
      for (auto& v : at[d].v) v *= X.theta;
    }
    applyop_4o(g, at, X.beta(0, 0, 0), x_u, b_u);
    std::vector<Fab> gp = make_vel(g, 0);
    compute_grad_4o(g, x_p, gp);
    for (int d = 0; d < g.dm; ++d) {
      Fab& B = b_u[d];
      for (int k = B.lo[2]; k <= B.hi[2]; ++k)
        for (int j = B.lo[1]; j <= B.hi[1]; ++j)
          for (int i = B.lo[0]; i <= B.hi[0]; ++i) B(i, j, k) += gp[d](i, j, k);
    }
    compute_div(g, x_u, b_p);
    for (int k = 0; k <= b_p.hi[2]; ++k)
      for (int j = 0; j <= b_p.hi[1]; ++j)
        for (int i = 0; i <= b_p.hi[0]; ++i) b_p(i, j, k) = -b_p(i, j, k);
    return;
  }
  Coef C;
  C.g = g;
  C.alpha = make_vel(g, 0);
  for (int d = 0; d < g.dm; ++d) {
    C.alpha[d].v = X.alpha[d].v;
    for (auto& v : C.alpha[d].v) v *= X.theta;
  }
  C.beta = &X.beta;
  C.gamma = &X.gamma;
  for (auto& e : X.bed) C.bed.push_back(&e);
  applyop(C, X.p.visc_type, x_u, b_u, 0);
  std::vector<Fab> gp = make_vel(g, 0);
  compute_grad(g, x_p, gp);
  for (int d = 0; d < g.dm; ++d) {
    Fab& B = b_u[d];
    for (int k = B.lo[2]; k <= B.hi[2]; ++k)
      for (int j = B.lo[1]; j <= B.hi[1]; ++j)
        for (int i = B.lo[0]; i <= B.hi[0]; ++i) B(i, j, k) += gp[d](i, j, k);
  }
  compute_div(g, x_u, b_p);
  for (int k = 0; k <= b_p.hi[2]; ++k)
    for (int j = 0; j <= b_p.hi[1]; ++j)
      for (int i = 0; i <= b_p.hi[0]; ++i) b_p(i, j, k) = -b_p(i, j, k);
}

// convert_to_homogeneous (convert_to_homogeneous.f90:19-106): phi = 0 with the inhomogeneous
// boundary values filled in (domainvel, periodic images, macvel), then b_u -= L phi (the
// stag_applyop operator with theta alpha) and b_p += D phi
void convert_to_homogeneous(Ctx& X, std::vector<Fab>& b_u, Fab& b_p, const VelBc& bc) {
  const Geom& g = X.g;
  std::vector<Fab> phi = make_vel(g, 1), Lphi = make_vel(g, 0);
  fill_vel_inh(phi, g, bc);
  Coef C;
  C.g = g;
  C.alpha = make_vel(g, 0);
  for (int d = 0; d < g.dm; ++d) {
    C.alpha[d].v = X.alpha[d].v;
    for (auto& v : C.alpha[d].v) v *= X.theta;
  }
  C.beta = &X.beta;
  C.gamma = &X.gamma;
  for (auto& e : X.bed) C.bed.push_back(&e);
  applyop(C, X.p.visc_type, phi, Lphi, 0);
  for (int d = 0; d < g.dm; ++d) {
    Fab& B = b_u[d];
    for (int k = B.lo[2]; k <= B.hi[2]; ++k)
      for (int j = B.lo[1]; j <= B.hi[1]; ++j)
        for (int i = B.lo[0]; i <= B.hi[0]; ++i) B(i, j, k) -= Lphi[d](i, j, k);
  }
  Fab dphi = make_cc(g, 0);
  compute_div(g, phi, dphi);
  for (int k = 0; k <= b_p.hi[2]; ++k)
    for (int j = 0; j <= b_p.hi[1]; ++j)
      for (int i = 0; i <= b_p.hi[0]; ++i) b_p(i, j, k) += dphi(i, j, k);
}

// sum_umac_press pressure part: mean over cells (norm_inner_product.f90:551-652)
double cc_sum(const Fab& p) {
  double s = 0.0;
  for (int k = 0; k <= p.hi[2]; ++k)
    for (int j = 0; j <= p.hi[1]; ++j)
      for (int i = 0; i <= p.hi[0]; ++i) s += p(i, j, k);
  return s;
}

// x_u -= w G p with w = 1/alpha (subtract_weighted_gradp, macproject.f90:318-362) or w = 1
static void sub_weighted_grad(Ctx& X, std::vector<Fab>& x_u, const Fab& p, bool unit_weight) {
  const Geom& g = X.g;
  std::vector<Fab> gp = make_vel(g, 0);
  compute_grad(g, p, gp);
  for (int d = 0; d < g.dm; ++d) {
    Fab& U = x_u[d];
    for (int k = gp[d].lo[2]; k <= gp[d].hi[2]; ++k)
      for (int j = gp[d].lo[1]; j <= gp[d].hi[1]; ++j)
        for (int i = gp[d].lo[0]; i <= gp[d].hi[0]; ++i)
          U(i, j, k) = unit_weight ? U(i, j, k) - gp[d](i, j, k)
                                   : U(i, j, k) - gp[d](i, j, k) * (1.0 / X.alpha[d](i, j, k));
  }
}

// x_p = -chi r + theta phi, chi = beta, 2 beta, 4/3 beta + gamma (the type-1 formula of
// apply_precon.f90:140-188 and simple_schur_complement, :461-515); theta phi only when
// add_phi (the MAC solve ran)
static void schur_simple(Ctx& X, const Fab& r, const Fab& phi, Fab& x_p, bool add_phi) {
  const int av = std::abs(X.p.visc_type);
  for (int k = 0; k <= x_p.hi[2]; ++k)
    for (int j = 0; j <= x_p.hi[1]; ++j)
      for (int i = 0; i <= x_p.hi[0]; ++i) {
        double v = -r(i, j, k);
        if (av == 1) v = v * X.beta(i, j, k);
        else if (av == 2) v = v * X.beta(i, j, k) * 2.0;
        else { const double t = v * X.gamma(i, j, k); v = v * X.beta(i, j, k) * (4.0 / 3.0) + t; }
        x_p(i, j, k) = add_phi ? v + phi(i, j, k) * X.theta : v;
      }
}

// L_alpha phi = D(alpha^-1 G phi) with Neumann walls: the level-0 MAC operator
static void mac_apply(Ctx& X, const Fab& phi_in, Fab& out);

// inverse_diag_lap (src_gmres/inverse_diag_lap.f90:17-250): Minv = 1 / the diagonal of A0 on
// every face, visc_type -1 / 1 / -2 / 2, each sum in the reference's order (in 2-D the node
// above precedes the node at the face, in 3-D the edge at the face precedes the one above)
static void inverse_diag_lap(Ctx& X, std::vector<Fab>& minv) {
  const Geom& g = X.g;
  const int dm = g.dm, vt = X.p.visc_type;
  const double b0 = X.beta(0, 0, 0);
  for (int a = 0; a < dm; ++a) {
    int tb[2];
    if (a == 0) { tb[0] = 1; tb[1] = 2; } else if (a == 1) { tb[0] = 0; tb[1] = 2; } else { tb[0] = 0; tb[1] = 1; }
    Fab& M = minv[a];
    for (int k = M.lo[2]; k <= M.hi[2]; ++k)
      for (int j = M.lo[1]; j <= M.hi[1]; ++j)
        for (int i = M.lo[0]; i <= M.hi[0]; ++i) {
          if (vt == 1) { M(i, j, k) = 1.0 / ((dm == 3 ? 6.0 : 4.0) * b0); continue; }
          if (vt == 2) { M(i, j, k) = 1.0 / ((dm == 3 ? 8.0 : 6.0) * b0); continue; }
          const int ea[3] = {a == 0, a == 1, a == 2};
          const double bI = X.beta(i, j, k), bM = X.beta(i - ea[0], j - ea[1], k - ea[2]);
          double sum = vt == -2 ? 2.0 * bI + 2.0 * bM : bI + bM;
          for (int t = 0; t < dm - 1; ++t) {
            const int b = tb[t];
            const int e = dm == 2 ? 0 : ((a + b) == 1 ? 0 : ((a + b) == 2 ? 1 : 2));
            const Fab& E = X.bed[e];
            const int eb[3] = {b == 0, b == 1, b == 2};
            const double e0 = E(i, j, k), e1 = E(i + eb[0], j + eb[1], k + eb[2]);
            sum = dm == 2 ? (sum + e1) + e0 : (sum + e0) + e1;
          }
          M(i, j, k) = 1.0 / sum;
        }
  }
}

// visc_schur_complement (apply_precon.f90:538-685): x_p = L_mu^-1 D Minv A0(beta) Minv G
// L_mu^-1 r + theta phi, L_mu = D(Minv G) the MAC operator with face coefficient Minv (zero on
// wall faces, Neumann), A0 applied with alpha = 0 and gamma = 0; gphi's ghosts are filled with
// the velocity rules (periodic images, reflection at walls -- the reference fills periodic
// ghosts only and reads unset ones next to a wall)
static void visc_schur(Ctx& X, const Fab& r, const Fab& phi, bool add_phi, Fab& x_p) {
  const Geom& g = X.g;
  const int dm = g.dm;
  std::vector<Fab> minv = make_vel(g, 0), bmu = make_vel(g, 0);
  inverse_diag_lap(X, minv);
  for (int d = 0; d < dm; ++d) {
    bmu[d].v = minv[d].v;
    for (int side = 0; side < 2; ++side) {
      if (!is_wall(side == 0 ? g.bclo[d] : g.bchi[d])) continue;
      const int idx = side == 0 ? 0 : g.n[d];
      Fab& B = bmu[d];
      for (int k = B.lo[2]; k <= B.hi[2]; ++k)
        for (int j = B.lo[1]; j <= B.hi[1]; ++j)
          for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
            int I[3] = {i, j, k};
            if (I[d] == idx) B(i, j, k) = 0.0;
          }
    }
  }
  Fab phimu = make_cc(g, 1);
  mac_mg(X, r, phimu, &bmu);
  std::vector<Fab> gphi = make_vel(g, 1), Lg = make_vel(g, 0);
  {
    std::vector<Fab> gp = make_vel(g, 0);
    compute_grad(g, phimu, gp);
    for (int d = 0; d < dm; ++d) {
      Fab& G = gphi[d];
      for (int k = gp[d].lo[2]; k <= gp[d].hi[2]; ++k)
        for (int j = gp[d].lo[1]; j <= gp[d].hi[1]; ++j)
          for (int i = gp[d].lo[0]; i <= gp[d].hi[0]; ++i) G(i, j, k) = gp[d](i, j, k) * minv[d](i, j, k);
    }
  }
  fill_vel(gphi, g);
  Coef C;
  C.g = g;
  C.alpha = make_vel(g, 0);
  Fab zero = make_cc(g, 1);
  C.beta = &X.beta;
  C.gamma = &zero;
  for (auto& e : X.bed) C.bed.push_back(&e);
  applyop(C, X.p.visc_type, gphi, Lg, 0);
  for (int d = 0; d < dm; ++d)
    for (size_t q = 0; q < Lg[d].v.size(); ++q) Lg[d].v[q] = Lg[d].v[q] * minv[d].v[q];
  std::vector<Fab> Lgg = make_vel(g, 1);
  for (int d = 0; d < dm; ++d) copy_valid(Lgg[d], Lg[d]);
  Fab r2 = make_cc(g, 0);
  compute_div(g, Lgg, r2);
  Fab phimu2 = make_cc(g, 1);
  mac_mg(X, r2, phimu2, &bmu);
  for (int k = 0; k <= x_p.hi[2]; ++k)
    for (int j = 0; j <= x_p.hi[1]; ++j)
      for (int i = 0; i <= x_p.hi[0]; ++i)
        x_p(i, j, k) = add_phi ? phimu2(i, j, k) + phi(i, j, k) * X.theta : phimu2(i, j, k);
}

// apply_precon (apply_precon.f90:29-460): precon_type +-1 .. +-5, visc_schur_approx 0 or 1
void apply_precon(Ctx& X, const std::vector<Fab>& b_u, const Fab& b_p, std::vector<Fab>& x_u, Fab& x_p) {
  const Geom& g = X.g;
  const int dm = g.dm;
  const int pt = X.p.precon_type, apt = std::abs(pt);
  Fab phi = make_cc(g, 1);
  Fab mac_rhs = make_cc(g, 0);
  for (int d = 0; d < dm; ++d) x_u[d].setval(0.0);
  auto div_plus_bp = [&]() {
    compute_div(g, x_u, mac_rhs);
    for (int k = 0; k <= mac_rhs.hi[2]; ++k)
      for (int j = 0; j <= mac_rhs.hi[1]; ++j)
        for (int i = 0; i <= mac_rhs.hi[0]; ++i) mac_rhs(i, j, k) += b_p(i, j, k);
  };
  // x_u = A^-1 (b_u - G x_p), V-cycles from the current x_u (zero for type 3, the first
  // solve's result for type 5)
  auto solve_u_minus_grad_xp = [&]() {
    std::vector<Fab> r = make_vel(g, 0);
    for (int d = 0; d < dm; ++d) copy_valid(r[d], b_u[d]);
    Fab xpg = make_cc(g, 1);
    copy_valid(xpg, x_p);
    fill_cc(xpg, g);
    sub_weighted_grad(X, r, xpg, true);
    stag_mg_solver(X, x_u, r);
  };
  if (apt == 1) {
    stag_mg_solver(X, x_u, b_u);
    div_plus_bp();
    mac_mg(X, mac_rhs, phi);
    sub_weighted_grad(X, x_u, phi, false);
    if (X.p.visc_schur_approx) {
      visc_schur(X, mac_rhs, phi, std::fabs(X.theta) > 0.0, x_p);
    } else if (pt == 1 || X.theta == 0.0) {
      schur_simple(X, mac_rhs, phi, x_p, true);
    } else {  // -1: -chi L_alpha phi in place of -chi mac_rhs (cc_applyop, :150-152)
      Fab lphi = make_cc(g, 0);
      mac_apply(X, phi, lphi);
      schur_simple(X, lphi, phi, x_p, true);
    }
  } else if (apt == 2 || apt == 5) {  // lower triangular (:196-276)
    stag_mg_solver(X, x_u, b_u);
    div_plus_bp();
    const bool mac = std::fabs(X.theta) > 0.0;
    if (mac) mac_mg(X, mac_rhs, phi);
    if (X.p.visc_schur_approx) visc_schur(X, mac_rhs, phi, mac, x_p);
    else schur_simple(X, mac_rhs, phi, x_p, mac);
    if (pt == -2 || pt == -5)
      for (auto& v : x_p.v) v = -v;
    if (apt == 5) solve_u_minus_grad_xp();
  } else if (apt == 3) {  // upper triangular (:278-344)
    for (int k = 0; k <= mac_rhs.hi[2]; ++k)
      for (int j = 0; j <= mac_rhs.hi[1]; ++j)
        for (int i = 0; i <= mac_rhs.hi[0]; ++i) mac_rhs(i, j, k) = pt == -3 ? -b_p(i, j, k) : b_p(i, j, k);
    const bool mac = std::fabs(X.theta) > 0.0;
    if (mac) mac_mg(X, mac_rhs, phi);
    if (X.p.visc_schur_approx) visc_schur(X, mac_rhs, phi, mac, x_p);
    else schur_simple(X, mac_rhs, phi, x_p, mac);
    solve_u_minus_grad_xp();
  } else {  // 4: block diagonal (:346-422)
    stag_mg_solver(X, x_u, b_u);
    const bool mac = std::fabs(X.theta) > 0.0;
    if (mac) {
      copy_valid(mac_rhs, b_p);
      mac_mg(X, mac_rhs, phi);
    }
    const int av = std::abs(X.p.visc_type);
    if (X.p.visc_schur_approx) {  // mac_rhs holds b_p only when the MAC solve ran (:358-365)
      visc_schur(X, mac_rhs, phi, mac, x_p);
    } else {
      for (int k = 0; k <= x_p.hi[2]; ++k)
        for (int j = 0; j <= x_p.hi[1]; ++j)
          for (int i = 0; i <= x_p.hi[0]; ++i) {
            const double tphi = mac ? phi(i, j, k) * X.theta : phi(i, j, k);
            double v = b_p(i, j, k) * X.beta(i, j, k) - tphi;
            if (av == 1 && pt == 4) v = v * -1.0;
            else if (av == 2 && pt == 4) v = v * -2.0;
            else if (av == 2 && pt == -4) v = v * 2.0;
            else if (av == 3 && pt == 4) v = v * (-4.0 / 3.0) + (b_p(i, j, k) * X.gamma(i, j, k)) * -1.0;
            else if (av == 3 && pt == -4) v = v * (4.0 / 3.0) + b_p(i, j, k) * X.gamma(i, j, k);
            x_p(i, j, k) = v;
          }
    }
  }
  const double mean = cc_sum(x_p) / (double)g.ncells();
  for (int k = 0; k <= x_p.hi[2]; ++k)
    for (int j = 0; j <= x_p.hi[1]; ++j)
      for (int i = 0; i <= x_p.hi[0]; ++i) x_p(i, j, k) -= mean;
  // theta == 0 with no no-slip wall would also remove the velocity mean (:441-458)
  if (X.theta == 0.0) {
    bool no_wall_is_no_slip = true;
    for (int d = 0; d < dm; ++d) if (is_noslip(g.bclo[d]) || is_noslip(g.bchi[d])) no_wall_is_no_slip = false;
    if (no_wall_is_no_slip) {
      for (int d = 0; d < dm; ++d) {
        if (!g.per(d)) continue;
        // face mean with 1/2 weights on the box-boundary faces (sum_umac_3d)
        double s = 0.0;
        Fab& U = x_u[d];
        for (int k = U.lo[2]; k <= U.hi[2]; ++k)
          for (int j = U.lo[1]; j <= U.hi[1]; ++j)
            for (int i = U.lo[0]; i <= U.hi[0]; ++i) {
              int I[3] = {i, j, k};
              double w = (I[d] == 0 || I[d] == g.n[d]) ? 0.5 : 1.0;
              s += w * U(i, j, k);
            }
        const double m = s / (double)g.ncells();
        for (int k = U.lo[2]; k <= U.hi[2]; ++k)
          for (int j = U.lo[1]; j <= U.hi[1]; ++j)
            for (int i = U.lo[0]; i <= U.hi[0]; ++i) U(i, j, k) -= m;
      }
    }
  }
}

// stag_inner_prod: 1/2 weight on the two box-boundary faces of each component's nodal
// direction (norm_inner_product.f90:183-313, single tile per box)
void stag_inner(const Geom& g, const std::vector<Fab>& a, const std::vector<Fab>& b, double* out) {
  for (int d = 0; d < g.dm; ++d) {
    const Fab& A = a[d];
    const Fab& B = b[d];
    const int kh = A.hi[2], jh = A.hi[1], ih = A.hi[0];
    double s = 0.0;
#pragma omp parallel for collapse(2) reduction(+ : s) schedule(static)
    for (int k = 0; k <= kh; ++k)
      for (int j = 0; j <= jh; ++j)
        for (int i = 0; i <= ih; ++i) {
          int I[3] = {i, j, k};
          const double w = (I[d] == 0 || I[d] == g.n[d]) ? 0.5 : 1.0;
          s += w * A(i, j, k) * B(i, j, k);
        }
    out[d] = s;
  }
}
double cc_inner(const Fab& a, const Fab& b) {
  double s = 0.0;
  const int kh = a.hi[2], jh = a.hi[1], ih = a.hi[0];
#pragma omp parallel for collapse(2) reduction(+ : s) schedule(static)
  for (int k = 0; k <= kh; ++k)
    for (int j = 0; j <= jh; ++j)
      for (int i = 0; i <= ih; ++i) s += a(i, j, k) * b(i, j, k);
  return s;
}

double norm2(const Geom& g, const std::vector<Fab>& u, const Fab& p, double pw) {
  double iu[3] = {0, 0, 0};
  stag_inner(g, u, u, iu);
  double nu = std::sqrt(iu[0] + iu[1] + iu[2]);
  double np = pw * std::sqrt(cc_inner(p, p));
  return std::sqrt(nu * nu + np * np);
}

// y += s*x over valid regions
void axpy_vel(std::vector<Fab>& y, double s, const std::vector<Fab>& x) {
  for (size_t d = 0; d < y.size(); ++d) {
    Fab& Y = y[d];
    const Fab& Xf = x[d];
    const int kh = Y.hi[2], jh = Y.hi[1], ih = Y.hi[0];
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k <= kh; ++k)
      for (int j = 0; j <= jh; ++j)
        for (int i = 0; i <= ih; ++i) Y(i, j, k) += s * Xf(i, j, k);
  }
}
void axpy_cc(Fab& y, double s, const Fab& x) {
  const int kh = y.hi[2], jh = y.hi[1], ih = y.hi[0];
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k <= kh; ++k)
    for (int j = 0; j <= jh; ++j)
      for (int i = 0; i <= ih; ++i) y(i, j, k) += s * x(i, j, k);
}

void rotmat(double a, double b, double& cs, double& sn) {
  if (b == 0.0) { cs = 1.0; sn = 0.0; }
  else if (std::fabs(b) > std::fabs(a)) {
    const double t = a / b;
    sn = 1.0 / std::sqrt(1.0 + t * t);
    cs = t * sn;
  } else {
    const double t = b / a;
    cs = 1.0 / std::sqrt(1.0 + t * t);
    sn = t * cs;
  }
}

// gmres (gmres.f90:32-621), scale_factor = 1
int gmres(Ctx& X, const std::vector<Fab>& b_u, const Fab& b_p, std::vector<Fab>& x_u, Fab& x_p, orc_stats* st) {
  const orc_params& p = X.p;
  const Geom& g = X.g;
  const int dm = g.dm;
  const int m = p.gmres_max_inner;
  const double pw = p.p_norm_weight;
  std::vector<Fab> r_u = make_vel(g, 1), w_u = make_vel(g, 0), tmp_u = make_vel(g, 1);
  Fab r_p = make_cc(g, 1), w_p = make_cc(g, 0), tmp_p = make_cc(g, 1);
  std::vector<std::vector<Fab>> V_u(m + 1);
  std::vector<Fab> V_p(m + 1);
  for (int i = 0; i <= m; ++i) { V_u[i] = make_vel(g, 0); V_p[i] = make_cc(g, 0); }
  std::vector<double> cs(m + 1), sn(m + 1), y(m + 1), s(m + 2);
  std::vector<double> H((size_t)(m + 2) * (m + 1), 0.0);
  auto Hm = [&](int r, int c) -> double& { return H[(size_t)r * (m + 1) + c]; };  // 1-based

  orc_stats S;
  std::memset(&S, 0, sizeof(S));
  auto precon = [&](const std::vector<Fab>& bu, const Fab& bp, std::vector<Fab>& xu, Fab& xp) {
    apply_precon(X, bu, bp, xu, xp);
    S.n_precon++;
  };
  auto matrix = [&](std::vector<Fab>& xu, Fab& xp, std::vector<Fab>& bu, Fab& bp) {
    apply_matrix(X, xu, xp, bu, bp);
    S.n_matrix++;
  };
  auto copy_vel = [&](std::vector<Fab>& dst, const std::vector<Fab>& src) {
    for (int d = 0; d < dm; ++d)
      for (int k = 0; k <= src[d].hi[2]; ++k)
        for (int j = 0; j <= src[d].hi[1]; ++j)
          for (int i = 0; i <= src[d].hi[0]; ++i) dst[d](i, j, k) = src[d](i, j, k);
  };
  auto copy_cc = [&](Fab& dst, const Fab& src) {
    for (int k = 0; k <= src.hi[2]; ++k)
      for (int j = 0; j <= src.hi[1]; ++j)
        for (int i = 0; i <= src.hi[0]; ++i) dst(i, j, k) = src(i, j, k);
  };

  precon(b_u, b_p, tmp_u, tmp_p);
  const double norm_pre_b = norm2(g, tmp_u, tmp_p, pw);
  const double norm_b = norm2(g, b_u, b_p, pw);
  S.norm_b = norm_b;
  S.norm_pre_b = norm_pre_b;
  if (norm_b <= p.gmres_abs_tol) {
    for (int d = 0; d < dm; ++d) x_u[d].setval(0.0);
    x_p.setval(0.0);
    S.converged = 1;
    if (st) *st = S;
    return 0;
  }
  int total_iter = 0, iter = 0;
  double norm_init_resid = 0.0, norm_resid = 0.0, norm_resid_stokes = 0.0;
  int nh = 0;
  while (true) {
    matrix(x_u, x_p, tmp_u, tmp_p);
    for (int d = 0; d < dm; ++d)
      for (int k = 0; k <= b_u[d].hi[2]; ++k)
        for (int j = 0; j <= b_u[d].hi[1]; ++j)
          for (int i = 0; i <= b_u[d].hi[0]; ++i) tmp_u[d](i, j, k) = -(tmp_u[d](i, j, k) - b_u[d](i, j, k));
    for (int k = 0; k <= b_p.hi[2]; ++k)
      for (int j = 0; j <= b_p.hi[1]; ++j)
        for (int i = 0; i <= b_p.hi[0]; ++i) tmp_p(i, j, k) = -(tmp_p(i, j, k) - b_p(i, j, k));
    norm_resid_stokes = norm2(g, tmp_u, tmp_p, pw);
    precon(tmp_u, tmp_p, r_u, r_p);
    norm_resid = norm2(g, r_u, r_p, pw);
    if (iter == 0) norm_init_resid = norm_resid;
    if (total_iter >= p.gmres_max_iter) break;
    if (total_iter >= p.gmres_min_iter) {
      if (norm_resid <= p.gmres_rel_tol * std::min(norm_pre_b, norm_init_resid)) { S.converged = 1; break; }
      if (norm_resid <= p.gmres_abs_tol) { S.converged = 1; break; }
    }
    if (iter >= p.gmres_max_outer) break;
    iter++;
    for (int d = 0; d < dm; ++d)
      for (int k = 0; k <= r_u[d].hi[2]; ++k)
        for (int j = 0; j <= r_u[d].hi[1]; ++j)
          for (int i = 0; i <= r_u[d].hi[0]; ++i) V_u[0][d](i, j, k) = r_u[d](i, j, k) / norm_resid;
    for (int k = 0; k <= r_p.hi[2]; ++k)
      for (int j = 0; j <= r_p.hi[1]; ++j)
        for (int i = 0; i <= r_p.hi[0]; ++i) V_p[0](i, j, k) = r_p(i, j, k) / norm_resid;
    std::fill(s.begin(), s.end(), 0.0);
    s[1] = norm_resid;
    int i_copy = 0;
    for (int i = 1; i <= m; ++i) {
      total_iter++;
      i_copy = i;
      copy_vel(r_u, V_u[i - 1]);
      copy_cc(r_p, V_p[i - 1]);
      matrix(r_u, r_p, tmp_u, tmp_p);
      precon(tmp_u, tmp_p, w_u, w_p);
      for (int k = 1; k <= i; ++k) {  // modified Gram-Schmidt
        double iu[3] = {0, 0, 0};
        stag_inner(g, w_u, V_u[k - 1], iu);
        const double ip = cc_inner(w_p, V_p[k - 1]);
        Hm(k, i) = (iu[0] + iu[1] + iu[2]) + (pw * pw) * ip;
        axpy_vel(w_u, -Hm(k, i), V_u[k - 1]);
        axpy_cc(w_p, -Hm(k, i), V_p[k - 1]);
      }
      Hm(i + 1, i) = norm2(g, w_u, w_p, pw);
      if (Hm(i + 1, i) == 0.0) { if (st) *st = S; return 2; }  // breakdown (gmres.f90:390-392)
      for (int d = 0; d < dm; ++d)
        for (int k = 0; k <= w_u[d].hi[2]; ++k)
          for (int j = 0; j <= w_u[d].hi[1]; ++j)
            for (int ii = 0; ii <= w_u[d].hi[0]; ++ii) V_u[i][d](ii, j, k) = w_u[d](ii, j, k) / Hm(i + 1, i);
      for (int k = 0; k <= w_p.hi[2]; ++k)
        for (int j = 0; j <= w_p.hi[1]; ++j)
          for (int ii = 0; ii <= w_p.hi[0]; ++ii) V_p[i](ii, j, k) = w_p(ii, j, k) / Hm(i + 1, i);
      // least_squares (gmres.f90:531-566)
      for (int k = 1; k <= i - 1; ++k) {
        const double t = cs[k] * Hm(k, i) + sn[k] * Hm(k + 1, i);
        Hm(k + 1, i) = -sn[k] * Hm(k, i) + cs[k] * Hm(k + 1, i);
        Hm(k, i) = t;
      }
      rotmat(Hm(i, i), Hm(i + 1, i), cs[i], sn[i]);
      const double t = cs[i] * s[i];
      s[i + 1] = -sn[i] * s[i];
      s[i] = t;
      Hm(i, i) = cs[i] * Hm(i, i) + sn[i] * Hm(i + 1, i);
      Hm(i + 1, i) = 0.0;
      const double est = std::fabs(s[i + 1]);
      if (nh < 256) S.resid_hist[nh++] = est;
      if (total_iter >= p.gmres_max_iter) break;
      if (total_iter >= p.gmres_min_iter) {
        if (est <= p.gmres_rel_tol * std::min(norm_pre_b, norm_init_resid) || est <= p.gmres_abs_tol) break;
      }
    }
    // SolveUTriangular(i_copy-1) then update_sol (gmres.f90:599-619,495-529)
    const int kk = i_copy - 1;
    y[kk + 1] = s[kk + 1] / Hm(kk + 1, kk + 1);
    for (int r = kk; r >= 1; --r) {
      double dot = 0.0;
      for (int c = r + 1; c <= kk + 1; ++c) dot += Hm(r, c) * y[c];
      y[r] = (s[r] - dot) / Hm(r, r);
    }
    for (int it = 1; it <= i_copy; ++it) {
      axpy_cc(x_p, y[it], V_p[it - 1]);
      axpy_vel(x_u, y[it], V_u[it - 1]);
    }
  }
  S.total_iter = total_iter;
  S.outer_iter = iter;
  S.norm_init_resid = norm_init_resid;
  S.norm_resid = norm_resid;
  S.norm_resid_stokes = norm_resid_stokes;
  if (st) *st = S;
  return 0;
}

// ---------------------------------------------------------------------------------------
// filter_stoch_m_flux (stochastic_m_fluxdiv.f90:604-778): along x, then y, then (3-D) z, each
// pass over the valid box (nodal extent included) from ghosts that multifab_fill_boundary
// refreshed before it; the taps and their order of summation are those of the reference:
// width 1: 1/2, 1/4; width 2: 5/8, 1/4, -1/16; width 4: 93/128, 7/32, -7/64, 1/32, -1/256
static double filter_taps(int w, double c, double m1, double p1, double m2, double p2, double m3, double p3,
                          double m4, double p4) {
  if (w == 1) return c / 2 + m1 / 4 + p1 / 4;
  if (w == 2) return 5 * c / 8 - m2 / 16 + m1 / 4 - p2 / 16 + p1 / 4;
  return 93 * c / 128 + 7 * m1 / 32 - 7 * m2 / 64 + m3 / 32 - m4 / 256 + 7 * p1 / 32 - 7 * p2 / 64 + p3 / 32 -
         p4 / 256;
}

static void filter_stoch(Fab& f, int w, const Geom& g) {
  for (int d = 0; d < g.dm; ++d) {
    fill_boundary(f, g);
    Fab t = f;
    for (int k = f.lo[2]; k <= f.hi[2]; ++k)
      for (int j = f.lo[1]; j <= f.hi[1]; ++j)
        for (int i = f.lo[0]; i <= f.hi[0]; ++i) {
          auto at = [&](int o) {
            int P[3] = {i, j, k};
            P[d] += o;
            return f.at(P);
          };
          const int r = w == 4 ? 4 : w;
          t(i, j, k) = filter_taps(w, at(0), at(-1), at(1), r >= 2 ? at(-2) : 0.0, r >= 2 ? at(2) : 0.0,
                                   r >= 4 ? at(-3) : 0.0, r >= 4 ? at(3) : 0.0, r >= 4 ? at(-4) : 0.0,
                                   r >= 4 ? at(4) : 0.0);
        }
    f = t;
  }
}

// stochastic_m_fluxdiv (stochastic_m_fluxdiv.f90:42-602), increment = .false.; with
// filtering_width > 0 the scaled stresses are filtered after their boundary sync (:172-175,
// :226-231), which this restatement supports on periodic boxes (the filter reads ghosts)
void stochastic_fluxdiv(Ctx& X, const double* const* Wcc, const double* const* Wed, const double* eta_cc,
                        const double* const* eta_ed, const double* T_cc, const double* const* T_ed, double dt,
                        double kT, double vcoef, double* const* fout) {
  const Geom& g = X.g;
  const int dm = g.dm;
  const double dz = dm == 3 ? g.dx : X.p.dz2d;
  const double variance = std::sqrt(vcoef * 2.0 * kT / (g.dx * g.dx * dz * dt));
  std::vector<Fab> fcc;
  Fab eta = make_cc(g, 0), T = make_cc(g, 0);
  load_valid(eta, eta_cc);
  if (T_cc) load_valid(T, T_cc); else T.setval(1.0);
  const int fw = X.p.filtering_width;
  for (int c = 0; c < dm; ++c) {
    Fab f = make_cc(g, std::max(1, fw));
    Fab w = make_cc(g, 0);
    load_valid(w, Wcc[c]);
    for (int k = 0; k <= w.hi[2]; ++k)
      for (int j = 0; j <= w.hi[1]; ++j)
        for (int i = 0; i <= w.hi[0]; ++i)
          f(i, j, k) = (variance * w(i, j, k)) * std::sqrt(eta(i, j, k) * T(i, j, k));
    fill_boundary(f, g);
    fcc.push_back(f);
  }
  const int ne = dm == 3 ? 3 : 1;
  std::vector<Fab> fed;  // ne arrays x 2 comps, stored as fed[2*e + c]
  for (int e = 0; e < ne; ++e) {
    Fab ee = make_edge(g, e, 0), tt = make_edge(g, e, 0);
    load_valid(ee, eta_ed[e]);
    if (T_ed) load_valid(tt, T_ed[e]); else tt.setval(1.0);
    for (int c = 0; c < 2; ++c) {
      Fab f = make_edge(g, e, 0);
      load_valid(f, Wed[2 * e + c]);
      for (size_t q = 0; q < f.v.size(); ++q) f.v[q] = (variance * f.v[q]) * std::sqrt(ee.v[q] * tt.v[q]);
      fed.push_back(f);
    }
  }
  // mflux_bc: x sqrt(2) on no-slip wall edges, 0 on slip wall edges (:395-451, 2-D :363-393)
  for (int d = 0; d < dm; ++d)
    for (int side = 0; side < 2; ++side) {
      const int code = side == 0 ? g.bclo[d] : g.bchi[d];
      if (!is_wall(code)) continue;
      const double fac = is_noslip(code) ? std::sqrt(2.0) : 0.0;
      const int idx = side == 0 ? 0 : g.n[d];
      for (int e = 0; e < ne; ++e) {
        bool touches = dm == 2 ? true : (e == 0 ? (d == 0 || d == 1) : e == 1 ? (d == 0 || d == 2) : (d == 1 || d == 2));
        if (!touches) continue;
        for (int c = 0; c < 2; ++c) {
          Fab& f = fed[2 * e + c];
          for (int k = f.lo[2]; k <= f.hi[2]; ++k)
            for (int j = f.lo[1]; j <= f.hi[1]; ++j)
              for (int i = f.lo[0]; i <= f.hi[0]; ++i) {
                int I[3] = {i, j, k};
                if (I[d] == idx) f(i, j, k) = fac * f(i, j, k);
              }
        }
      }
    }
  // multifab_internal_sync: duplicated periodic nodes take the value at index 0
  for (auto& f : fed)
    for (int d = 0; d < dm; ++d) {
      if (!g.per(d) || f.hi[d] != g.n[d]) continue;
      for (int k = f.lo[2]; k <= f.hi[2]; ++k)
        for (int j = f.lo[1]; j <= f.hi[1]; ++j)
          for (int i = f.lo[0]; i <= f.hi[0]; ++i) {
            int I[3] = {i, j, k};
            if (I[d] == g.n[d]) { int S0[3] = {i, j, k}; S0[d] = 0; f(i, j, k) = f.at(S0); }
          }
    }
  if (fw > 0) {
    for (size_t r = 0; r < fed.size(); ++r) {
      Fab& f = fed[r];
      Fab h = make_edge(g, (int)r / 2, fw);
      std::vector<double> v(f.v.size());
      store_valid(f, v.data());
      load_valid(h, v.data());
      filter_stoch(h, fw, g);
      store_valid(h, v.data());
      load_valid(f, v.data());
    }
    for (auto& f : fcc) {
      filter_stoch(f, fw, g);
      fill_boundary(f, g);  // the first ghost is read by the divergence
    }
  }
  // divergence (stoch_m_force_3d :508-600, 2-D :453-506)
  const double dxinv = 1.0 / g.dx;
  std::vector<Fab> out = make_vel(g, 0);
  for (int a = 0; a < dm; ++a) {
    Fab& F = out[a];
    for (int k = F.lo[2]; k <= F.hi[2]; ++k)
      for (int j = F.lo[1]; j <= F.hi[1]; ++j)
        for (int i = F.lo[0]; i <= F.hi[0]; ++i) {
          int I[3] = {i, j, k};
          // contributions in direction order x, y, z
          double v = 0.0;
          for (int b = 0; b < dm; ++b) {
            int P[3] = {i, j, k};
            if (b == a) {
              int M[3] = {i, j, k}; M[a] -= 1;
              v += (fcc[a].at(I) - fcc[a].at(M)) * dxinv;
            } else {
              P[b] += 1;
              const int e = dm == 2 ? 0 : edge_id(a, b);
              const int c = a < b ? 0 : 1;
              v += (fed[2 * e + c].at(P) - fed[2 * e + c].at(I)) * dxinv;
            }
          }
          F(i, j, k) = v;
        }
    domainvel(F, a, g);
    store_valid(F, fout[a]);
  }
}

// average_cc_to_face: 2-point mean; wall faces take the ghost value (convert_stag.f90:293-393)
void avg_face(const Geom& g, const Fab& cc, std::vector<Fab>& f) {
  for (int a = 0; a < g.dm; ++a) {
    Fab& F = f[a];
    for (int k = F.lo[2]; k <= F.hi[2]; ++k)
      for (int j = F.lo[1]; j <= F.hi[1]; ++j)
        for (int i = F.lo[0]; i <= F.hi[0]; ++i) {
          int I[3] = {i, j, k}, M[3] = {i, j, k};
          M[a] -= 1;
          double v = 0.5 * (cc.at(I) + cc.at(M));
          if (I[a] == 0 && is_wall(g.bclo[a])) v = cc.at(M);
          if (I[a] == g.n[a] && is_wall(g.bchi[a])) v = cc.at(I);
          F(i, j, k) = v;
        }
  }
}

// average_cc_to_edge (3-D, convert_stag.f90:681-862) / average_cc_to_node (2-D, :462-611):
// 4-point mean; edges on a wall take the mean of the two ghost cells beside them
void avg_edge(const Geom& g, const Fab& cc, std::vector<Fab>& ed) {
  const int ne = g.dm == 3 ? 3 : 1;
  for (int e = 0; e < ne; ++e) {
    int p0, p1;  // the two nodal directions of this edge type
    if (g.dm == 2 || e == 0) { p0 = 0; p1 = 1; }
    else if (e == 1) { p0 = 0; p1 = 2; }
    else { p0 = 1; p1 = 2; }
    Fab& E = ed[e];
    for (int k = E.lo[2]; k <= E.hi[2]; ++k)
      for (int j = E.lo[1]; j <= E.hi[1]; ++j)
        for (int i = E.lo[0]; i <= E.hi[0]; ++i) {
          int I[3] = {i, j, k};
          int A[3] = {i, j, k}, B[3] = {i, j, k}, Cc[3] = {i, j, k};
          A[p0] -= 1; B[p1] -= 1; Cc[p0] -= 1; Cc[p1] -= 1;
          double v = 0.25 * (cc.at(I) + cc.at(A) + cc.at(B) + cc.at(Cc));
          // walls in the two nodal directions, processed in direction order (later wins)
          for (int w = 0; w < 2; ++w) {
            const int d = w == 0 ? p0 : p1;
            const int o = w == 0 ? p1 : p0;
            for (int side = 0; side < 2; ++side) {
              const int code = side == 0 ? g.bclo[d] : g.bchi[d];
              if (!is_wall(code)) continue;
              const int idx = side == 0 ? 0 : g.n[d];
              if (I[d] != idx) continue;
              int G0[3] = {i, j, k};
              G0[d] = side == 0 ? -1 : g.n[d];
              int G1[3] = {G0[0], G0[1], G0[2]};
              G1[o] -= 1;
              v = 0.5 * (cc.at(G0) + cc.at(G1));
            }
          }
          E(i, j, k) = v;
        }
  }
}

void load_vel(std::vector<Fab>& u, const double* const* src) {
  for (size_t d = 0; d < u.size(); ++d) load_valid(u[d], src[d]);
}
void store_vel(const std::vector<Fab>& u, double* const* dst) {
  for (size_t d = 0; d < u.size(); ++d) store_valid(u[d], dst[d]);
}

Ctx* C_(void* c) { return reinterpret_cast<Ctx*>(c); }

}  // namespace

// ===========================================================================================
extern "C" {

void orc_default_params(orc_params* p, int dm, int nx, int ny, int nz, double dx) {
  std::memset(p, 0, sizeof(*p));
  p->dm = dm;
  p->n[0] = nx; p->n[1] = ny; p->n[2] = dm == 3 ? nz : 1;
  p->dx = dx;
  p->dz2d = 1.0;
  for (int d = 0; d < 3; ++d) { p->bc_lo[d] = -1; p->bc_hi[d] = -1; }
  p->visc_type = 1;
  p->stag_nu_down = 2; p->stag_nu_up = 2; p->stag_nu_bottom = 8;
  p->stag_minwidth = 2; p->stag_max_vcycles = 1; p->stag_smoother = 1;
  p->stag_omega = 1.0; p->stag_rel_tol = 1e-9;
  p->mac_nu_down = 2; p->mac_nu_up = 2; p->mac_nu_bottom = 8; p->mac_minwidth = 2; p->mac_max_vcycles = 1;
  p->gmres_max_inner = 5; p->gmres_max_outer = 20; p->gmres_max_iter = 100; p->gmres_min_iter = 1;
  p->gmres_rel_tol = 1e-9; p->gmres_abs_tol = 0.0; p->p_norm_weight = 1.0;
  p->precon_type = 1;
  p->omp_threads = 0;
  p->gmres_spatial_order = 2;
  p->visc_schur_approx = 0;
  p->filtering_width = 0;
}

void* orc_create(const orc_params* p) {
  Ctx* X = new Ctx;
  X->p = *p;
  X->g.dm = p->dm;
  for (int d = 0; d < 3; ++d) {
    X->g.n[d] = (d < p->dm) ? p->n[d] : 1;
    X->g.bclo[d] = p->bc_lo[d];
    X->g.bchi[d] = p->bc_hi[d];
  }
  X->g.dx = p->dx;
#ifdef _OPENMP
  if (p->omp_threads > 0) omp_set_num_threads(p->omp_threads);
#endif
  return X;
}

void orc_destroy(void* c) { delete C_(c); }

int orc_set_coeffs(void* c, const double* const* alpha_fc, const double* beta_cc, const double* const* beta_ed,
                   const double* gamma_cc, double theta) {
  Ctx& X = *C_(c);
  const Geom& g = X.g;
  X.alpha = make_vel(g, 0);
  load_vel(X.alpha, alpha_fc);
  X.beta = make_cc(g, 1);
  X.gamma = make_cc(g, 1);
  load_valid(X.beta, beta_cc);
  load_valid(X.gamma, gamma_cc);
  fill_cc(X.beta, g);
  fill_cc(X.gamma, g);
  X.bed.clear();
  for (int e = 0; e < (g.dm == 3 ? 3 : 1); ++e) {
    X.bed.push_back(make_edge(g, e, 0));
    load_valid(X.bed[e], beta_ed[e]);
  }
  X.theta = theta;
  X.have_coef = true;
  return 0;
}

int orc_stag_applyop(void* c, const double* const* phi, double* const* Lphi, int color) {
  Ctx& X = *C_(c);
  const Geom& g = X.g;
  std::vector<Fab> u = make_vel(g, 1), L = make_vel(g, 0);
  load_vel(u, phi);
  fill_vel(u, g);
  Coef C;
  C.g = g;
  C.alpha = make_vel(g, 0);
  for (int d = 0; d < g.dm; ++d) {
    C.alpha[d].v = X.alpha[d].v;
    for (auto& v : C.alpha[d].v) v *= X.theta;
  }
  C.beta = &X.beta; C.gamma = &X.gamma;
  for (auto& e : X.bed) C.bed.push_back(&e);
  // untouched (uncoloured) faces are returned as the input Lphi values
  load_vel(L, reinterpret_cast<const double* const*>(Lphi));
  applyop(C, X.p.visc_type, u, L, color);
  store_vel(L, Lphi);
  return 0;
}

// Level-0 coefficients of the context in multigrid form (alpha scaled by theta)
static Coef level0_coef(Ctx& X) {
  const Geom& g = X.g;
  Coef C;
  C.g = g;
  C.alpha = make_vel(g, 0);
  for (int d = 0; d < g.dm; ++d) {
    C.alpha[d].v = X.alpha[d].v;
    for (auto& v : C.alpha[d].v) v *= X.theta;
  }
  C.beta = &X.beta;
  C.gamma = &X.gamma;
  for (auto& e : X.bed) C.bed.push_back(&e);
  return C;
}

// One colour of the staggered multigrid smoother on the context's grid, as the V-cycle runs it
// (stag_mg_solver.f90:363-396): L phi on the colour, stag_mg_update on the colour, ghost fill.
int orc_stag_smooth_colour(void* c, double* const* phi, const double* const* rhs, int color) {
  Ctx& X = *C_(c);
  const Geom& g = X.g;
  Coef C = level0_coef(X);
  std::vector<Fab> u = make_vel(g, 1), r = make_vel(g, 0), L = make_vel(g, 1);
  load_vel(u, reinterpret_cast<const double* const*>(phi));
  load_vel(r, rhs);
  fill_vel(u, g);
  applyop(C, X.p.visc_type, u, L, color);
  mg_update(C, X.p.visc_type, u, r, L, color, X.p.stag_omega);
  fill_vel(u, g);
  store_vel(u, phi);
  return 0;
}

// Residual restriction of the V-cycle (stag_mg_solver.f90:966-1170, 12/6-point) from the
// context's grid to the grid coarsened by 2: ghosts of the fine field filled by the boundary
// rules first, wall faces of the result zeroed as the V-cycle does.  simple != 0: the 4/2-point
// mean used for alpha.
int orc_stag_restriction(void* c, const double* const* fine, double* const* coarse, int simple) {
  Ctx& X = *C_(c);
  const Geom& g = X.g;
  const Geom gc = g.coarsen(2);
  std::vector<Fab> f = make_vel(g, 1), cf = make_vel(gc, 0);
  load_vel(f, fine);
  fill_vel(f, g);
  stag_restriction(gc, f, cf, simple != 0);
  if (!simple)
    for (int d = 0; d < g.dm; ++d) domainvel(cf[d], d, gc);
  store_vel(cf, coarse);
  return 0;
}

// Prolongation-add of the V-cycle (stag_mg_solver.f90:1365-1628): fine += P coarse, with the
// coarse field on the grid coarsened by 2 and its ghosts filled by the boundary rules.
int orc_stag_prolongation(void* c, const double* const* coarse, double* const* fine) {
  Ctx& X = *C_(c);
  const Geom& g = X.g;
  const Geom gc = g.coarsen(2);
  std::vector<Fab> f = make_vel(g, 1), cf = make_vel(gc, 1);
  load_vel(cf, coarse);
  fill_vel(cf, gc);
  load_vel(f, reinterpret_cast<const double* const*>(fine));
  stag_prolongation(g, f, cf);
  store_vel(f, fine);
  return 0;
}

int orc_apply_matrix(void* c, const double* const* x_u, const double* x_p, double* const* b_u, double* b_p) {
  Ctx& X = *C_(c);
  const Geom& g = X.g;
  std::vector<Fab> u = make_vel(g, 1), bu = make_vel(g, 0);
  Fab p = make_cc(g, 1), bp = make_cc(g, 0);
  load_vel(u, x_u);
  load_valid(p, x_p);
  apply_matrix(X, u, p, bu, bp);
  store_vel(bu, b_u);
  store_valid(bp, b_p);
  return 0;
}

static VelBc load_velbc(const Geom& g, const double* const* vbn, const double* const* vbt) {
  VelBc bc;
  bc.n = make_vel(g, 0);
  load_vel(bc.n, vbn);
  const int nt = g.dm == 3 ? 6 : 2;
  for (int q = 0; q < nt; ++q) {
    int d = g.dm == 2 ? q : q / 2;
    int b = g.dm == 2 ? 1 - q : (q % 2 == 0 ? (d == 0 ? 1 : 0) : (d == 2 ? 1 : 2));
    Fab f = make_edge(g, g.dm == 2 ? 0 : edge_id(d, b), 0);
    load_valid(f, vbt[q]);
    bc.t.push_back(f);
  }
  return bc;
}

int orc_apply_matrix_bc(void* c, const double* const* x_u, const double* x_p, double* const* b_u, double* b_p,
                        const double* const* vel_bc_n, const double* const* vel_bc_t) {
  Ctx& X = *C_(c);
  const Geom& g = X.g;
  if (X.p.gmres_spatial_order != 2) return 1;
  VelBc bc = load_velbc(g, vel_bc_n, vel_bc_t);
  std::vector<Fab> u = make_vel(g, 1), bu = make_vel(g, 0);
  Fab p = make_cc(g, 1), bp = make_cc(g, 0);
  load_vel(u, x_u);
  load_valid(p, x_p);
  apply_matrix(X, u, p, bu, bp, &bc);
  store_vel(bu, b_u);
  store_valid(bp, b_p);
  return 0;
}

int orc_convert_to_homogeneous(void* c, double* const* b_u, double* b_p, const double* const* vel_bc_n,
                               const double* const* vel_bc_t) {
  Ctx& X = *C_(c);
  const Geom& g = X.g;
  VelBc bc = load_velbc(g, vel_bc_n, vel_bc_t);
  std::vector<Fab> bu = make_vel(g, 0);
  Fab bp = make_cc(g, 0);
  load_vel(bu, reinterpret_cast<const double* const*>(b_u));
  load_valid(bp, b_p);
  convert_to_homogeneous(X, bu, bp, bc);
  store_vel(bu, b_u);
  store_valid(bp, b_p);
  return 0;
}

int orc_apply_precon(void* c, const double* const* b_u, const double* b_p, double* const* x_u, double* x_p) {
  Ctx& X = *C_(c);
  const Geom& g = X.g;
  std::vector<Fab> bu = make_vel(g, 0), xu = make_vel(g, 1);
  Fab bp = make_cc(g, 0), xp = make_cc(g, 1);
  load_vel(bu, b_u);
  load_valid(bp, b_p);
  apply_precon(X, bu, bp, xu, xp);
  store_vel(xu, x_u);
  store_valid(xp, x_p);
  return 0;
}

int orc_stag_mg_solve(void* c, const double* const* rhs, double* const* phi) {
  Ctx& X = *C_(c);
  const Geom& g = X.g;
  std::vector<Fab> r = make_vel(g, 0), u = make_vel(g, 1);
  load_vel(r, rhs);
  load_vel(u, reinterpret_cast<const double* const*>(phi));
  stag_mg_solver(X, u, r);
  store_vel(u, phi);
  return 0;
}

int orc_mac_mg_solve(void* c, const double* rhs, double* phi) {
  Ctx& X = *C_(c);
  Fab r = make_cc(X.g, 0), p = make_cc(X.g, 1);
  load_valid(r, rhs);
  mac_mg(X, r, p);
  store_valid(p, phi);
  return 0;
}

int orc_gmres(void* c, const double* const* b_u, const double* b_p, double* const* x_u, double* x_p, orc_stats* st) {
  Ctx& X = *C_(c);
  const Geom& g = X.g;
  std::vector<Fab> bu = make_vel(g, 0), xu = make_vel(g, 1);
  Fab bp = make_cc(g, 0), xp = make_cc(g, 1);
  load_vel(bu, b_u);
  load_valid(bp, b_p);
  load_vel(xu, reinterpret_cast<const double* const*>(x_u));
  load_valid(xp, x_p);
  int rc = gmres(X, bu, bp, xu, xp, st);
  store_vel(xu, x_u);
  store_valid(xp, x_p);
  return rc;
}

// mk_advective_m_fluxdiv (src_lowMach/mk_advective_m_fluxdiv.f90:184-345 in 3-D, :93-182 in
// 2-D): centred momentum advection -div(m u) on every valid face, with the momentum m = rho u
// formed on the valid faces (convert_m_to_umac) and its ghosts filled like a velocity; the
// update on wall faces is zero (multifab_physbc_domainvel, :81-86).  Flux of m_a through the
// d-faces of its control volume: f(c) = 1/4 (m_a(c - e_d) + m_a(c)) (u_d(c - e_a) + u_d(c)), and
// m_update_a = -(sum over d = x, y, z of (f(c + e_d) - f(c)) / dx), or m_update_a minus that sum
// with increment.
int orc_advective_m_fluxdiv(void* c, const double* const* u_in, const double* const* rho_in, double* const* out,
                            int increment) {
  Ctx& X = *C_(c);
  const Geom& g = X.g;
  const int dm = g.dm;
  std::vector<Fab> u = make_vel(g, 1), m = make_vel(g, 1), A = make_vel(g, 0), rho = make_vel(g, 0);
  load_vel(u, u_in);
  load_vel(rho, rho_in);
  fill_vel(u, g);
  for (int a = 0; a < dm; ++a) {
    Fab& M = m[a];
    for (int k = rho[a].lo[2]; k <= rho[a].hi[2]; ++k)
      for (int j = rho[a].lo[1]; j <= rho[a].hi[1]; ++j)
        for (int i = rho[a].lo[0]; i <= rho[a].hi[0]; ++i) M(i, j, k) = rho[a](i, j, k) * u[a](i, j, k);
  }
  fill_vel(m, g);
  if (increment) load_vel(A, reinterpret_cast<const double* const*>(out));
  const double dxinv = 1.0 / g.dx;
  for (int a = 0; a < dm; ++a) {
    const int ea[3] = {a == 0, a == 1, a == 2};
    Fab& R = A[a];
    for (int k = R.lo[2]; k <= R.hi[2]; ++k)
      for (int j = R.lo[1]; j <= R.hi[1]; ++j)
        for (int i = R.lo[0]; i <= R.hi[0]; ++i) {
          const int I[3] = {i, j, k};
          if ((I[a] == 0 && is_wall(g.bclo[a])) || (I[a] == g.n[a] && is_wall(g.bchi[a]))) {
            R(i, j, k) = 0.0;
            continue;
          }
          double s = 0.0;
          for (int d = 0; d < dm; ++d) {
            const int ed[3] = {d == 0, d == 1, d == 2};
            auto f = [&](int ci, int cj, int ck) {
              return 0.25 * (m[a](ci - ed[0], cj - ed[1], ck - ed[2]) + m[a](ci, cj, ck)) *
                     (u[d](ci - ea[0], cj - ea[1], ck - ea[2]) + u[d](ci, cj, ck));
            };
            const double t = (f(i + ed[0], j + ed[1], k + ed[2]) - f(i, j, k)) * dxinv;
            s = d == 0 ? t : s + t;
          }
          R(i, j, k) = increment ? R(i, j, k) - s : -s;
        }
  }
  store_vel(A, out);
  return 0;
}

// diffusive_m_fluxdiv (src_lowMach/diffusive_m_fluxdiv.f90:17-90): stag_applyop with alpha = 0
// gives -A0 u; the result is its negative, or with increment it is subtracted from out
int orc_diffusive_m_fluxdiv(void* c, const double* const* u_in, double* const* out, int increment) {
  Ctx& X = *C_(c);
  const Geom& g = X.g;
  std::vector<Fab> u = make_vel(g, 1), L = make_vel(g, 0), R = make_vel(g, 0);
  load_vel(u, u_in);
  fill_vel(u, g);
  Coef C;
  C.g = g;
  C.alpha = make_vel(g, 0);  // zero
  C.beta = &X.beta; C.gamma = &X.gamma;
  for (auto& e : X.bed) C.bed.push_back(&e);
  applyop(C, X.p.visc_type, u, L, 0);
  if (increment) load_vel(R, reinterpret_cast<const double* const*>(out));
  for (int d = 0; d < g.dm; ++d)
    for (size_t e = 0; e < R[d].v.size(); ++e) R[d].v[e] = increment ? R[d].v[e] - L[d].v[e] : -L[d].v[e];
  store_vel(R, out);
  return 0;
}

// the absolute GMRES tolerance of later solves (gmres_abs_tol of probin_gmres.f90, which the
// time steps overwrite between their two solves)
void orc_set_gmres_abs_tol(void* c, double tol) { C_(c)->p.gmres_abs_tol = tol; }

int orc_compute_div(void* c, const double* const* u, double* div) {
  Ctx& X = *C_(c);
  std::vector<Fab> uu = make_vel(X.g, 1);
  load_vel(uu, u);
  Fab d = make_cc(X.g, 0);
  compute_div(X.g, uu, d);
  store_valid(d, div);
  return 0;
}

int orc_compute_grad(void* c, const double* p, double* const* gout) {
  Ctx& X = *C_(c);
  Fab pp = make_cc(X.g, 1);
  load_valid(pp, p);
  fill_cc(pp, X.g);
  std::vector<Fab> gp = make_vel(X.g, 0);
  compute_grad(X.g, pp, gp);
  store_vel(gp, gout);
  return 0;
}

int orc_inner_prod(void* c, const double* const* u1, const double* p1, const double* const* u2, const double* p2,
                   double* out_u, double* out_p) {
  Ctx& X = *C_(c);
  std::vector<Fab> a = make_vel(X.g, 0), b = make_vel(X.g, 0);
  load_vel(a, u1);
  load_vel(b, u2);
  stag_inner(X.g, a, b, out_u);
  Fab pa = make_cc(X.g, 0), pb = make_cc(X.g, 0);
  load_valid(pa, p1);
  load_valid(pb, p2);
  *out_p = cc_inner(pa, pb);
  return 0;
}

int orc_stochastic_m_fluxdiv(void* c, const double* const* W_cc, const double* const* W_ed, const double* eta_cc,
                             const double* const* eta_ed, const double* T_cc, const double* const* T_ed, double dt,
                             double kT, double variance_coef, double* const* f) {
  stochastic_fluxdiv(*C_(c), W_cc, W_ed, eta_cc, eta_ed, T_cc, T_ed, dt, kT, variance_coef, f);
  return 0;
}

int orc_average_cc_to_face(void* c, const double* cc_g1, double* const* face) {
  Ctx& X = *C_(c);
  Fab cc = make_cc(X.g, 1);
  std::memcpy(cc.v.data(), cc_g1, cc.v.size() * sizeof(double));
  std::vector<Fab> f = make_vel(X.g, 0);
  avg_face(X.g, cc, f);
  store_vel(f, face);
  return 0;
}

int orc_average_cc_to_edge(void* c, const double* cc_g1, double* const* edge) {
  Ctx& X = *C_(c);
  Fab cc = make_cc(X.g, 1);
  std::memcpy(cc.v.data(), cc_g1, cc.v.size() * sizeof(double));
  std::vector<Fab> e;
  for (int q = 0; q < (X.g.dm == 3 ? 3 : 1); ++q) e.push_back(make_edge(X.g, q, 0));
  avg_edge(X.g, cc, e);
  for (size_t q = 0; q < e.size(); ++q) store_valid(e[q], edge[q]);
  return 0;
}

int orc_nlevs_mg(void* c, int minwidth) { return nlevs_mg(C_(c)->g, minwidth); }

}  // extern "C"
