// stencils.h -- per-point device code of the staggered Stokes solver (gfx950).
//
// Every kernel is a small functor whose operator()(t) handles one face / cell / edge; launch.h
// maps a flat index space onto a 1-D grid.  Fields live on their valid region only (layout in
// lowmachfhd_amd/grid.py); out-of-range neighbours are resolved in-register:
//   * periodic direction: the periodic image (a nodal index n+1 maps to 1, -1 to n-1, so the
//     duplicated face n stays its own degree of freedom exactly as in FBoxLib);
//   * wall, transverse velocity: -interior (no-slip) / +interior (slip), the ghost rule of
//     multifab_physbc_macvel (src_common/multifab_physbc_stag.f90:1276-1645);
//   * wall, cell-centred scalar: the adjacent interior value (FOEXTRAP);
//   * wall-normal faces on a wall are held at zero (multifab_physbc_domainvel).
// Reference formulas: src_gmres/stag_applyop.f90 (operator, 2-D :212-515, 3-D :517-1075),
// src_gmres/stag_mg_solver.f90 (update :1632-2458, transfers :868-1628).
#pragma once
#include <hip/hip_runtime.h>

#define HD __host__ __device__ __forceinline__

namespace lmfhd {

struct Geo {
  int dm;
  int n[3];          // cells (n[2] = 1 in 2-D)
  int per[3];        // periodic flags
  int wlo[3], whi[3];// wall flags
  int slo[3], shi[3];// transverse-velocity ghost sign at a wall: -1 no-slip, +1 slip
};

HD int imin(int x, int y) { return x < y ? x : y; }
HD int fext(const Geo& g, int a, int d) { return g.n[d] + (d == a ? 1 : 0); }
HD long fsize(const Geo& g, int a) { return (long)fext(g, a, 0) * fext(g, a, 1) * fext(g, a, 2); }
HD long csize(const Geo& g) { return (long)g.n[0] * g.n[1] * g.n[2]; }
HD int enodal(int dm, int e, int d) {
  if (dm == 2) return d < 2;
  return e == 0 ? (d < 2) : (e == 1 ? (d != 1) : (d > 0));
}
HD int eext(const Geo& g, int e, int d) { return g.n[d] + enodal(g.dm, e, d); }
HD long esize(const Geo& g, int e) { return (long)eext(g, e, 0) * eext(g, e, 1) * eext(g, e, 2); }
HD int edge_id(int dm, int a, int b) {
  if (dm == 2) return 0;
  const int s = a + b;
  return s == 1 ? 0 : (s == 2 ? 1 : 2);
}
// array offsets in 32-bit arithmetic (every array holds fewer than 2^31 points)
HD int fidx(const Geo& g, int a, int i, int j, int k) { return i + fext(g, a, 0) * (j + fext(g, a, 1) * k); }
HD int cidx(const Geo& g, int i, int j, int k) { return i + g.n[0] * (j + g.n[1] * k); }
HD int eidx(const Geo& g, int e, int i, int j, int k) { return i + eext(g, e, 0) * (j + eext(g, e, 1) * k); }
HD bool wall_face(const Geo& g, int a, int ia) { return (ia == 0 && g.wlo[a]) || (ia == g.n[a] && g.whi[a]); }

// Parity-split storage of a multigrid level's face arrays (phi, rhs, residual, alpha): face
// (i,j,k) of component a lives at q H + (i >> 1) + hx (j + ny k), q = (i + j + k) & 1, with
// hx = (nx + 1) / 2 and H = hx ny nz over component a's extents.  The faces of one parity are
// then contiguous, so a colour sweep reads and writes whole cache lines instead of every other
// value of each line.  Colour item u of parity q (decode_color) is the face at q H + u.
HD unsigned shx(const Geo& g, int a) { return (unsigned)((fext(g, a, 0) + 1) >> 1); }
HD unsigned sH(const Geo& g, int a) { return shx(g, a) * (unsigned)(fext(g, a, 1) * fext(g, a, 2)); }
HD unsigned sidx(const Geo& g, int a, int i, int j, int k) {
  return (unsigned)((i + j + k) & 1) * sH(g, a) + (unsigned)(i >> 1) + shx(g, a) * (unsigned)(j + fext(g, a, 1) * k);
}
// index of face (i,j,k) in natural (sp = 0) or parity-split (sp = 1) storage
HD unsigned lidx(const Geo& g, int a, int i, int j, int k, int sp) {
  return sp ? sidx(g, a, i, j, k) : (unsigned)fidx(g, a, i, j, k);
}
// elements a face array of component a occupies in that storage
inline long fstore(const Geo& g, int a, int sp) { return sp ? 2L * sH(g, a) : fsize(g, a); }

// p[idx] through a 32-bit unsigned byte offset (every array here is smaller than 4 GiB, which
// ctx_build checks): with a wave-uniform base the access takes the scalar-base + 32-bit
// vector-offset form instead of a 64-bit address formed per lane
template <typename T>
HD T ld(const T* __restrict__ p, unsigned idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + idx * (unsigned)sizeof(T));
}
template <typename T>
HD void st(T* p, unsigned idx, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + idx * (unsigned)sizeof(T)) = v;
}

// velocity component a at face (i,j,k), possibly outside the valid region.  A wall ghost of
// the transverse velocity reflects onto the face itself; `alt`, when given, is read there
// instead of p.  WD: the leading directions resolved (3: all; 1: x only, for callers whose y
// and z indices are known to be inside the box -- rows away from the y and z boundaries).
// sp: p in parity-split storage.
template <typename T, int WD = 3>
HD T vel_at(const T* __restrict__ p, const Geo& g, int a, int i, int j, int k, const T* __restrict__ alt = nullptr,
            int sp = 0) {
  // Written with selects and unconditional loads of the geometry: with branches, every
  // boundary point ran a chain of divergent blocks, each waiting on its own scalar load.
  int I[3] = {i, j, k};
  int sg = 1;
  bool refl = false;
#pragma unroll
  for (int d = 0; d < WD; ++d) {
    const int n = g.n[d];
    const int v = I[d];
    const bool lo = v < 0;
    if (d == a) {
      const bool hi = v > n;
      I[d] = v + (lo ? n : 0) - (hi ? n : 0);
    } else {
      const bool hi = v >= n;
      const int per = g.per[d];  // 0 or 1: blended arithmetically, not branched on
      const int slo = g.slo[d], shi = g.shi[d];
      const int vp = v + (lo ? n : 0) - (hi ? n : 0);
      const int vw = lo ? 0 : (hi ? n - 1 : v);
      I[d] = vw + per * (vp - vw);
      const int sw = lo ? slo : (hi ? shi : 1);
      sg *= sw + per * (1 - sw);
      refl = refl || ((1 - per) * (int)(lo || hi) != 0);
    }
  }
  const T* q = (refl && alt) ? alt : p;
  const T v = ld(q, lidx(g, a, I[0], I[1], I[2], sp));
  return sg < 0 ? -v : v;  // the ghost sign: the same bits as a product with -1 / +1
}

// cell-centred value at (i,j,k), periodic image or adjacent interior cell at a wall (WD as in
// vel_at)
template <typename T, int WD = 3>
HD T cc_at(const T* __restrict__ p, const Geo& g, int i, int j, int k) {
  int I[3] = {i, j, k};
#pragma unroll
  for (int d = 0; d < WD; ++d) {
    const int n = g.n[d];
    const int v = I[d];
    const bool lo = v < 0, hi = v >= n;
    const int vp = v + (lo ? n : 0) - (hi ? n : 0);
    const int vw = lo ? 0 : (hi ? n - 1 : v);
    I[d] = vw + g.per[d] * (vp - vw);
  }
  return ld(p, (unsigned)cidx(g, I[0], I[1], I[2]));
}

template <typename T>
struct Coefs {
  const T* alpha[3];  // face
  const T* beta;      // cell
  const T* gamma;     // cell
  const T* bed[3];    // edges (3-D) / node (2-D)
  T dxsqinv;
  T theta;            // alpha is multiplied by theta on the fly (1 on multigrid levels)
};

// f(0), f(1) or f(2) for component a: every call is inlined with its component known, so
// the code under it has no branches on the component and loads its operands together
template <int DM, class F>
HD void by_comp(int a, const F& f) {
  if (a == 0) f(0);
  else if (DM == 2 || a == 1) f(1);
  else f(2);
}

HD void transverse(int a, int& b0, int& b1) {
  if (a == 0) { b0 = 1; b1 = 2; }
  else if (a == 1) { b0 = 0; b1 = 2; }
  else { b0 = 0; b1 = 1; }
}

// (theta alpha - A0) phi at face (i,j,k) of component a, and the diagonal coefficient.
// Generic in the roles (normal a, transverse b): reproduces the term order of the x, y, z
// blocks of stag_applyop_3d for visc_type -1/-2/-3 and the constant forms 1/2/3.
// `ctr`, when given, holds the value of component a at the face itself (and so at its wall
// reflections); the neighbours are read from phi[a].
// Component a's values for the stencil come through an accessor: GlobalAcc reads global
// memory (neighbours through the wrap / reflect rules; `ctr` as above), BcGhostAcc the lift
// field of inhomogeneous walls.
template <typename T, int WD = 3>
struct GlobalAcc {
  const T* p;
  const T* ctr;
  const Geo* g;
  int a;
  int sp = 0;  // component a's arrays (and alpha) in parity-split storage
  HD unsigned idx(int i, int j, int k) const { return lidx(*g, a, i, j, k, sp); }
  HD T c(unsigned oa) const { return ctr ? ld(ctr, oa) : ld(p, oa); }
  HD T n(int i, int j, int k) const { return vel_at<T, WD>(p, *g, a, i, j, k, ctr, sp); }
};
// WD as in vel_at: the wrap / reflect rules applied to the cell and cross-term neighbours
template <typename T, int VT, int DM, class Acc, int WD = 3>
HD T lphi_acc(const Geo& g, const Coefs<T>& C, T* const* phi, int a, int i, int j, int k, T* diag_out,
              const Acc& A) {
  const int ea0 = a == 0, ea1 = a == 1, ea2 = a == 2;
  int tb[2];
  transverse(a, tb[0], tb[1]);
  const unsigned oa = A.idx(i, j, k);
  const T alpha = ld(C.alpha[a], oa) * C.theta;
  const T pa = A.c(oa);
  const T pa_p = A.n(i + ea0, j + ea1, k + ea2);
  const T pa_m = A.n(i - ea0, j - ea1, k - ea2);
  if constexpr (VT < 0) {
    const T bI = cc_at<T, WD>(C.beta, g, i, j, k);
    const T bM = cc_at<T, WD>(C.beta, g, i - ea0, j - ea1, k - ea2);
    T gI = T(0), gM = T(0), nI, nM;
    if constexpr (VT == -3) {
      gI = cc_at<T, WD>(C.gamma, g, i, j, k);
      gM = cc_at<T, WD>(C.gamma, g, i - ea0, j - ea1, k - ea2);
      nI = T(4.0 / 3.0) * bI + gI;
      nM = T(4.0 / 3.0) * bM + gM;
    } else if constexpr (VT == -2) {
      nI = T(2) * bI;
      nM = T(2) * bM;
    } else {
      nI = bI;
      nM = bM;
    }
    T dsum = nI + nM;
    T offs = pa_p * nI + pa_m * nM;
    T cross = T(0);
#pragma unroll
    for (int t = 0; t < DM - 1; ++t) {
      const int b = tb[t];
      const int eb0 = b == 0, eb1 = b == 1, eb2 = b == 2;
      const int e = edge_id(DM, a, b);
      const T* E = C.bed[e];
      const T eI = ld(E, (unsigned)eidx(g, e, i, j, k));
      const T eP = ld(E, (unsigned)eidx(g, e, i + eb0, j + eb1, k + eb2));
      dsum += eI + eP;
      offs += A.n(i + eb0, j + eb1, k + eb2) * eP + A.n(i - eb0, j - eb1, k - eb2) * eI;
      if constexpr (VT != -1) {
        const T pbP = vel_at<T, WD>(phi[b], g, b, i + eb0, j + eb1, k + eb2, nullptr, A.sp);
        const T pbI = vel_at<T, WD>(phi[b], g, b, i, j, k, nullptr, A.sp);
        const T pbMP = vel_at<T, WD>(phi[b], g, b, i - ea0 + eb0, j - ea1 + eb1, k - ea2 + eb2, nullptr, A.sp);
        const T pbM = vel_at<T, WD>(phi[b], g, b, i - ea0, j - ea1, k - ea2, nullptr, A.sp);
        if constexpr (VT == -2) {
          cross += pbP * eP - pbI * eI - pbMP * eP + pbM * eI;
        } else {
          const T tI = -T(2.0 / 3.0) * bI + gI, tM = -T(2.0 / 3.0) * bM + gM;
          cross += pbP * (eP + tI) - pbI * (eI + tI) - pbMP * (eP + tM) + pbM * (eI + tM);
        }
      }
    }
    const T diag = alpha + dsum * C.dxsqinv;
    if (diag_out) *diag_out = diag;
    return pa * diag - (offs + cross) * C.dxsqinv;
  } else {
    const T b0 = C.beta[0];
    const T nsum = pa_p + pa_m;
    T tsum = T(0), cross = T(0);
#pragma unroll
    for (int t = 0; t < DM - 1; ++t) {
      const int b = tb[t];
      const int eb0 = b == 0, eb1 = b == 1, eb2 = b == 2;
      tsum += A.n(i + eb0, j + eb1, k + eb2) + A.n(i - eb0, j - eb1, k - eb2);
      if constexpr (VT != 1) {
        cross += vel_at<T, WD>(phi[b], g, b, i + eb0, j + eb1, k + eb2, nullptr, A.sp) -
                 vel_at<T, WD>(phi[b], g, b, i, j, k, nullptr, A.sp) -
                 vel_at<T, WD>(phi[b], g, b, i - ea0 + eb0, j - ea1 + eb1, k - ea2 + eb2, nullptr, A.sp) +
                 vel_at<T, WD>(phi[b], g, b, i - ea0, j - ea1, k - ea2, nullptr, A.sp);
      }
    }
    T diag, L;
    if constexpr (VT == 1) {
      diag = alpha + T(2 * DM) * b0 * C.dxsqinv;
      L = pa * diag - (nsum + tsum) * b0 * C.dxsqinv;
    } else if constexpr (VT == 2) {
      diag = alpha + T(2 * DM + 2) * b0 * C.dxsqinv;
      L = pa * diag - (T(2) * nsum + tsum + cross) * b0 * C.dxsqinv;
    } else {
      const T c0 = C.gamma[0];
      const T num = DM == 3 ? T(20) : T(14);
      diag = alpha + (num * b0 * T(1.0 / 3.0) + T(2) * c0) * C.dxsqinv;
      L = pa * diag - (nsum * (T(4.0 / 3.0) * b0 + c0) + tsum * b0 + cross * (T(1.0 / 3.0) * b0 + c0)) * C.dxsqinv;
    }
    if (diag_out) *diag_out = diag;
    return L;
  }
}

template <typename T, int VT, int DM, int WD = 3>
HD T lphi(const Geo& g, const Coefs<T>& C, T* const* phi, int a, int i, int j, int k, T* diag_out,
          const T* __restrict__ ctr = nullptr, int sp = 0) {
  const GlobalAcc<T, WD> A{phi[a], ctr, &g, a, sp};
  return lphi_acc<T, VT, DM, GlobalAcc<T, WD>, WD>(g, C, phi, a, i, j, k, diag_out, A);
}

// diagonal coefficient only (stag_mg_update "fac")
template <typename T, int VT, int DM>
HD T diag_at(const Geo& g, const Coefs<T>& C, int a, int i, int j, int k) {
  const int ea0 = a == 0, ea1 = a == 1, ea2 = a == 2;
  const T alpha = ld(C.alpha[a], (unsigned)fidx(g, a, i, j, k)) * C.theta;
  if constexpr (VT > 0) {
    const T b0 = C.beta[0];
    if constexpr (VT == 1) return alpha + T(2 * DM) * b0 * C.dxsqinv;
    if constexpr (VT == 2) return alpha + T(2 * DM + 2) * b0 * C.dxsqinv;
    const T num = DM == 3 ? T(20) : T(14);
    return alpha + (num * b0 * T(1.0 / 3.0) + T(2) * C.gamma[0]) * C.dxsqinv;
  } else {
    const T bI = cc_at(C.beta, g, i, j, k), bM = cc_at(C.beta, g, i - ea0, j - ea1, k - ea2);
    T nI, nM;
    if constexpr (VT == -3) {
      nI = T(4.0 / 3.0) * bI + cc_at(C.gamma, g, i, j, k);
      nM = T(4.0 / 3.0) * bM + cc_at(C.gamma, g, i - ea0, j - ea1, k - ea2);
    } else if constexpr (VT == -2) {
      nI = T(2) * bI;
      nM = T(2) * bM;
    } else {
      nI = bI;
      nM = bM;
    }
    T dsum = nI + nM;
    int tb[2];
    transverse(a, tb[0], tb[1]);
#pragma unroll
    for (int t = 0; t < DM - 1; ++t) {
      const int b = tb[t];
      const int e = edge_id(DM, a, b);
      dsum += ld(C.bed[e], (unsigned)eidx(g, e, i, j, k)) +
              ld(C.bed[e], (unsigned)eidx(g, e, i + (b == 0), j + (b == 1), k + (b == 2)));
    }
    return alpha + dsum * C.dxsqinv;
  }
}

// ---------------------------------------------------------------------------------------
// flat-index decoders.  Division by a launch-invariant extent uses the multiply-high method of
// Granlund & Montgomery ("Division by invariant integers using multiplication", 1994, fig 4.1)
// for numerators below 2^31: m = floor(2^32 (2^l - d)/d) + 1, q = (mulhi(m, u) + u) >> l.
struct FastDiv {
  unsigned d, m, l;
};
inline FastDiv make_fastdiv(unsigned d) {
  unsigned l = 0;
  while ((1u << l) < d) ++l;
  const unsigned long long m = ((1ull << 32) * ((1ull << l) - d)) / d + 1ull;
  return FastDiv{d, (unsigned)m, l};
}
HD unsigned fdiv(const FastDiv& f, unsigned u) {
  const unsigned t = (unsigned)(((unsigned long long)f.m * u) >> 32);
  return (t + u) >> f.l;
}
// extents of a launch: rows of nx (nh = (nx+1)/2 for one colour) by ny by nz
struct Dec {
  FastDiv x, y;
  int nx, ny;
};
inline Dec make_dec(int nx, int ny) { return Dec{make_fastdiv((unsigned)nx), make_fastdiv((unsigned)ny), nx, ny}; }
inline Dec make_dec_color(int nx, int ny) {
  return Dec{make_fastdiv((unsigned)((nx + 1) / 2)), make_fastdiv((unsigned)ny), nx, ny};
}
HD void decode3(long t, const Dec& D, int& i, int& j, int& k) {
  const unsigned u = (unsigned)t;
  const unsigned r = fdiv(D.x, u);
  i = (int)(u - r * D.x.d);
  const unsigned q = fdiv(D.y, r);
  j = (int)(r - q * D.y.d);
  k = (int)q;
}
HD bool decode_color(long t, const Dec& D, int par, int& i, int& j, int& k) {
  const unsigned u = (unsigned)t;
  const unsigned r = fdiv(D.x, u);
  const int ih = (int)(u - r * D.x.d);
  const unsigned q = fdiv(D.y, r);
  j = (int)(r - q * D.y.d);
  k = (int)q;
  i = 2 * ih + ((par + j + k) & 1);
  return i < D.nx;
}
// L phi (mode 0) or rhs - L phi (mode 1) on every face of component a; zero on wall faces
template <typename T, int VT, int DM>
struct ApplyFn {
  Geo g;
  Coefs<T> C;
  T* phi[3];
  const T* rhs;
  T* out;
  int a;
  Dec dc;
  int mode;
  // the component through by_comp: with a runtime component the coarse-level kernels indexed
  // the functor's arrays dynamically and kept it in scratch memory
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) {
      const int I[3] = {i, j, k};
      if (wall_face(g, ac, I[ac])) { out[t] = T(0); return; }
      const T L = lphi<T, VT, DM>(g, C, phi, ac, i, j, k, nullptr);
      out[t] = mode == 0 ? L : rhs[t] - L;
    });
  }
};

// face (i,j,k) exists for component a
HD bool face_in(const Geo& g, int a, int i, int j, int k) {
  return i < fext(g, a, 0) && j < fext(g, a, 1) && k < fext(g, a, 2);
}

// all components at once: L phi (mode 0) or rhs - L phi (mode 1), zero on wall faces.  One
// thread per (i,j) column of the nodal (n+1)-box marches over a run of k, so every array is
// streamed once per launch and the k-1/k+1 planes are re-read from cache.
template <typename T, int VT, int DM>
struct ApplyAllFn {
  Geo g;
  Coefs<T> C;
  T* phi[3];
  const T* rhs[3];
  T* out[3];
  FastDiv cx;  // divides by nx + 1
  int mode;
  HD void operator()(int col, int k0, int k1) const {
    const int j = (int)fdiv(cx, (unsigned)col);
    const int i = col - j * (int)cx.d;
    for (int k = k0; k < k1; ++k) {
      // every component at the nearest face of its box (loads in bounds), then the stores
      T v[DM];
      unsigned o[DM];
#pragma unroll
      for (int a = 0; a < DM; ++a) {
        const int ic = imin(i, fext(g, a, 0) - 1), jc = imin(j, fext(g, a, 1) - 1), kc = imin(k, fext(g, a, 2) - 1);
        const int I[3] = {ic, jc, kc};
        o[a] = (unsigned)fidx(g, a, ic, jc, kc);
        const T r = mode == 0 ? T(0) : ld(rhs[a], o[a]);
        const T L = lphi<T, VT, DM>(g, C, phi, a, ic, jc, kc, nullptr);
        v[a] = wall_face(g, a, I[a]) ? T(0) : (mode == 0 ? L : r - L);
      }
#pragma unroll
      for (int a = 0; a < DM; ++a)
        if (face_in(g, a, i, j, k)) st(out[a], o[a], v[a]);
    }
  }
};

// ApplyAllFn with one row per wave (launch_rows): wave w walks the points (i, j, k) of one row
// of the nodal box with i < n_x (lanes along x, chunks of 64), j <= n_y, k <= n_z, so j and k
// are wave-uniform; the x faces i = n_x follow as tail waves with lanes along y, one k plane
// per ct chunks.  Every point gets the arithmetic of ApplyAllFn, so the results are the same
// bit for bit.
template <typename T, int VT, int DM>
struct ApplyRowsFn {
  Geo g;
  Coefs<T> C;
  T* phi[3];
  const T* rhs[3];
  T* out[3];
  int mode;
  int wt;       // first tail wave
  FastDiv cw;   // chunks per row
  FastDiv nyr;  // rows along y: n_y + 1
  FastDiv ct;   // tail chunks per k plane
  HD void operator()(int w, int lane) const {
    if (w < wt) {
      const int r = (int)fdiv(cw, (unsigned)w);
      const int c = w - r * (int)cw.d;
      const int k = (int)fdiv(nyr, (unsigned)r);
      const int j = r - k * (int)nyr.d;
      const int i = 64 * c + lane;
      if (i >= g.n[0]) return;
      // rows whose y and z neighbours all lie inside the box resolve x only (a uniform branch)
      if (j >= 1 && j <= g.n[1] - 2 && (DM == 2 || (k >= 1 && k <= g.n[2] - 2))) point<1>(i, j, k);
      else point<3>(i, j, k);
    } else {
      const int u = w - wt;
      const int k = (int)fdiv(ct, (unsigned)u);
      const int c = u - k * (int)ct.d;
      const int j = 64 * c + lane;
      if (j >= g.n[1]) return;
      const int i = g.n[0];
      const unsigned o = (unsigned)fidx(g, 0, i, j, k);
      const T rr = mode == 0 ? T(0) : ld(rhs[0], o);
      const T L = lphi<T, VT, DM>(g, C, phi, 0, i, j, k, nullptr);
      st(out[0], o, wall_face(g, 0, i) ? T(0) : (mode == 0 ? L : rr - L));
    }
  }
  template <int WD>
  HD void point(int i, int j, int k) const {
    T v[DM];
    unsigned o[DM];
#pragma unroll
    for (int a = 0; a < DM; ++a) {
      const int jc = imin(j, fext(g, a, 1) - 1), kc = imin(k, fext(g, a, 2) - 1);
      const int I[3] = {i, jc, kc};
      o[a] = (unsigned)fidx(g, a, i, jc, kc);
      const T rr = mode == 0 ? T(0) : ld(rhs[a], o[a]);
      const T L = lphi<T, VT, DM, WD>(g, C, phi, a, i, jc, kc, nullptr);
      v[a] = wall_face(g, a, I[a]) ? T(0) : (mode == 0 ? L : rr - L);
    }
#pragma unroll
    for (int a = 0; a < DM; ++a)
      if (face_in(g, a, i, j, k)) st(out[a], o[a], v[a]);
  }
};

// L phi on the faces of one colour only (phase 1 of the two-phase update)
template <typename T, int VT, int DM>
struct ApplyColorFn {
  Geo g;
  Coefs<T> C;
  T* phi[3];
  T* out;
  int a;
  Dec dc;
  int par;
  HD void operator()(long t) const {
    int i, j, k;
    if (!decode_color(t, dc, par, i, j, k)) return;
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) {
      const int I[3] = {i, j, k};
      const long o = fidx(g, ac, i, j, k);
      if (wall_face(g, ac, I[ac])) { out[o] = T(0); return; }
      out[o] = lphi<T, VT, DM>(g, C, phi, ac, i, j, k, nullptr);
    });
  }
};

// phi += omega (rhs - Lphi)/diag on one colour with Lphi from a buffer (phase 2)
template <typename T, int VT, int DM>
struct UpdateColorFn {
  Geo g;
  Coefs<T> C;
  T* phi_a;
  const T* rhs_a;
  const T* L_a;
  T omega;
  int a;
  Dec dc;
  int par;
  HD void operator()(long t) const {
    int i, j, k;
    if (!decode_color(t, dc, par, i, j, k)) return;
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) {
      const int I[3] = {i, j, k};
      if (wall_face(g, ac, I[ac])) return;
      const long o = fidx(g, ac, i, j, k);
      const T diag = diag_at<T, VT, DM>(g, C, ac, i, j, k);
      phi_a[o] = phi_a[o] + omega * (rhs_a[o] - L_a[o]) / diag;
    });
  }
};

// fused in-place colour Gauss-Seidel step: valid when the red-black colouring is consistent
// across the periodic wrap (even sizes), where no face of the colour reads another
template <typename T, int VT, int DM>
struct GsColorFn {
  Geo g;
  Coefs<T> C;
  T* phi[3];
  const T* rhs_a;
  T omega;
  int a;
  Dec dc;
  int par;
  HD void operator()(long t) const {
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) {
      int i, j, k;
      if (!decode_color(t, dc, par, i, j, k)) return;
      const int I[3] = {i, j, k};
      if (wall_face(g, ac, I[ac])) return;
      const unsigned o = (unsigned)fidx(g, ac, i, j, k);
      const T old = ld(phi[ac], o), r = ld(rhs_a, o);  // loaded with the stencil's operands
      T diag;
      const T L = lphi<T, VT, DM>(g, C, phi, ac, i, j, k, &diag);
      st(phi[ac], o, old + omega * (r - L) / diag);
    });
  }
};

// One parity of every component in one launch, for |visc_type| = 1: there the operator of
// component a reads component a only (no cross terms), so the colours of different components
// commute and colour 2a+1 of every a, then colour 2a+2 of every a, equals the reference's order
// 1, 2, ..., 2 dm bit for bit. With ilv = 0, items [0, n0) are component 0's, [n0, n01)
// component 1's, the rest component 2's (each the colour items of GsColorFn). With ilv = 1 the
// components alternate by blocks of 256 items: block b holds items 256 (b / DM) + [0, 256) of
// component b % DM, so the three components' faces of one region are smoothed by neighbouring
// workgroups, which share the cell and edge coefficients they read through one XCD's L2
// (launch DM * ceil(max nc / 256) * 256 items).  SP = 1: phi, rhs and alpha in parity-split
// storage, where item u of parity par is the face at par H + u.
template <typename T, int VT, int DM, int SP = 0>
struct GsColorAllFn {
  static_assert(VT == 1 || VT == -1, "components are coupled for |visc_type| > 1");
  Geo g;
  Coefs<T> C;
  T* phi[3];
  const T* rhs[3];
  T omega;
  Dec dc[3];
  long n0, n01;
  unsigned nc[3];  // colour items per component
  int ilv;
  int par;
  HD void operator()(long t) const {
    int a;
    long u;
    if (ilv) {
      const unsigned blk = (unsigned)(t >> 8), q = blk / DM;
      a = (int)(blk - q * DM);
      const unsigned uu = (q << 8) + ((unsigned)t & 255u);
      if (uu >= (a == 0 ? nc[0] : (a == 1 ? nc[1] : nc[2]))) return;
      u = uu;
    } else {  // [0, n0) component 0, [n0, n01) component 1, the rest component 2; a range may
              // end in padding (nc: the component's items)
      a = t < n0 ? 0 : (t < n01 ? 1 : 2);
      u = t - (a == 0 ? 0 : (a == 1 ? n0 : n01));
    }
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) {
      // with the literal component: a runtime index into nc kept a local functor in scratch
      if (u >= (long)nc[ac]) return;
      int i, j, k;
      if (!decode_color(u, dc[ac], par, i, j, k)) return;
      const int I[3] = {i, j, k};
      if (wall_face(g, ac, I[ac])) return;
      const unsigned o = SP ? (unsigned)par * sH(g, ac) + (unsigned)u : (unsigned)fidx(g, ac, i, j, k);
      const T old = ld(phi[ac], o), r = ld(rhs[ac], o);
      T diag;
      const T L = lphi<T, VT, DM>(g, C, phi, ac, i, j, k, &diag, nullptr, SP);
      st(phi[ac], o, old + omega * (r - L) / diag);
    });
  }
};

// The multigrid residual res = rhs - L phi (mode 1) or L phi (mode 0) on a level in parity-split
// storage, one face per item in storage order: block b holds items 256 (b / DM) + [0, 256) of
// component b % DM, item v = q H + u being colour item u of parity q (slack items of odd extents
// do nothing).  Every array of the face is then read and written along whole lines, where the
// column launch's lanes would alternate between the two halves.  Each face gets ApplyAllFn's
// arithmetic; wall faces get zero.
template <typename T, int VT, int DM>
struct ResidSplitFn {
  static_assert(VT == 1 || VT == -1, "split storage only for |visc_type| 1");
  Geo g;
  Coefs<T> C;
  T* phi[3];
  const T* rhs[3];
  T* out[3];
  Dec dc[3];      // colour decoders (hx, ny)
  unsigned H[3];  // items per parity
  int mode;
  HD void operator()(long t) const {
    const unsigned blk = (unsigned)(t >> 8), qb = blk / DM;
    const int a = (int)(blk - qb * DM);
    const unsigned v = (qb << 8) + ((unsigned)t & 255u);
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) {
      const unsigned h = H[ac];
      if (v >= 2 * h) return;
      const int par = v >= h ? 1 : 0;
      int i, j, k;
      if (!decode_color(v - (unsigned)par * h, dc[ac], par, i, j, k)) return;
      const int I[3] = {i, j, k};
      const T r = mode == 0 ? T(0) : ld(rhs[ac], v);
      const T L = lphi<T, VT, DM>(g, C, phi, ac, i, j, k, nullptr, nullptr, 1);
      st(out[ac], v, wall_face(g, ac, I[ac]) ? T(0) : (mode == 0 ? L : r - L));
    });
  }
};

// Split-storage levels, one thread per position (i,j,k) of the nodal box handling the face of
// every component at that position: GS = 1, the colour Gauss-Seidel step of parity par (the
// faces of one position share its parity; the same update as GsColorAllFn); GS = 0, the
// residual rhs - L phi (mode 1) or L phi (mode 0) of ResidSplitFn, over both parities.  Items
// follow the split order of the nodal box (item u of parity q: decode_color over (n_x + 1,
// n_y + 1)), so every component's faces are read and written along whole lines, and the three
// share the cell and edge coefficients and the index arithmetic of their position.  Where the
// position lies outside a component's box the face nearest to it is evaluated (loads in bounds)
// and the value discarded.  Measured at 128^3 in FP32: 20 us against 26 us per Gauss-Seidel
// launch, 35 against 50 us per residual; in FP64 the one-face-per-thread kernels are as fast.
template <typename T, int VT, int DM, int GS>
struct SplitPos3Fn {
  static_assert(VT == 1 || VT == -1, "split storage only for |visc_type| 1");
  Geo g;
  Coefs<T> C;
  T* phi[3];
  const T* rhs[3];
  T* out[3];
  T omega;
  Dec dn;       // colour decoder of the nodal box (n_x + 1, n_y + 1)
  unsigned Hn;  // positions per parity
  int par;      // GS: the parity swept
  int mode;     // residual: 1 rhs - L phi, 0 L phi
  HD void operator()(long t) const {
    int q = par;
    unsigned u = (unsigned)t;
    if (!GS) {
      q = u >= Hn ? 1 : 0;
      u -= (unsigned)q * Hn;
    }
    int i, j, k;
    if (!decode_color(u, dn, q, i, j, k)) return;
    T v[DM];
    unsigned o[DM];
    bool put[DM];
#pragma unroll
    for (int a = 0; a < DM; ++a) {
      const int ic = imin(i, fext(g, a, 0) - 1), jc = imin(j, fext(g, a, 1) - 1), kc = imin(k, fext(g, a, 2) - 1);
      const int I[3] = {ic, jc, kc};
      o[a] = sidx(g, a, ic, jc, kc);
      put[a] = face_in(g, a, i, j, k);
      if (GS) {
        const T old = ld(phi[a], o[a]), r = ld(rhs[a], o[a]);
        T diag;
        const T L = lphi<T, VT, DM>(g, C, phi, a, ic, jc, kc, &diag, nullptr, 1);
        v[a] = old + omega * (r - L) / diag;
        put[a] = put[a] && !wall_face(g, a, I[a]);
      } else {
        const T r = mode == 0 ? T(0) : ld(rhs[a], o[a]);
        const T L = lphi<T, VT, DM>(g, C, phi, a, ic, jc, kc, nullptr, nullptr, 1);
        v[a] = wall_face(g, a, I[a]) ? T(0) : (mode == 0 ? L : r - L);
      }
    }
#pragma unroll
    for (int a = 0; a < DM; ++a)
      if (put[a]) st(GS ? phi[a] : out[a], o[a], v[a]);
  }
};

// GsColorAllFn with one row per wave (launch_rows): wave w walks the colour items of one (j, k)
// row of one component, lanes along x, so j and k are wave-uniform.  Waves [0, wb[0]) are
// component 0's rows (x faces 0 .. n_x - 1), [wb[0], wb[1]) component 1's, [wb[1], wb[2])
// component 2's (none in 2-D), each row cw[a] chunks of 64 items; the x faces i = n_x follow
// as tail waves with lanes along y, one k plane per ct chunks.  Every face gets the arithmetic
// of GsColorAllFn, so the sweep is the same bit for bit.
template <typename T, int VT, int DM>
struct GsRowsFn {
  static_assert(VT == 1 || VT == -1, "components are coupled for |visc_type| > 1");
  Geo g;
  Coefs<T> C;
  T* phi[3];
  const T* rhs[3];
  T omega;
  int par;
  int wb[3];
  FastDiv cw[3];  // chunks per row
  FastDiv ny[3];  // rows along y
  FastDiv ct;     // tail chunks per k plane
  HD void operator()(int w, int lane) const {
    if (w < wb[0]) row<0>(w, lane);
    else if (w < wb[1]) row<1>(w - wb[0], lane);
    else if (DM == 3 && w < wb[2]) row<DM - 1>(w - wb[1], lane);
    else tail(w - wb[2], lane);
  }
  template <int AC>
  HD void row(int w, int lane) const {
    const int r = (int)fdiv(cw[AC], (unsigned)w);
    const int c = w - r * (int)cw[AC].d;
    const int k = (int)fdiv(ny[AC], (unsigned)r);
    const int j = r - k * (int)ny[AC].d;
    const int i = 2 * (64 * c + lane) + ((par + j + k) & 1);
    if (i >= fext(g, AC, 0) - (AC == 0 ? 1 : 0)) return;
    // rows whose y and z neighbours all lie inside the box resolve x only (a uniform branch)
    if (j >= 1 && j <= fext(g, AC, 1) - 2 && (DM == 2 || (k >= 1 && k <= fext(g, AC, 2) - 2))) at<AC, 1>(i, j, k);
    else at<AC, 3>(i, j, k);
  }
  HD void tail(int w, int lane) const {
    const int k = (int)fdiv(ct, (unsigned)w);
    const int c = w - k * (int)ct.d;
    const int i = g.n[0];
    const int j = 2 * (64 * c + lane) + ((par + i + k) & 1);
    if (j >= g.n[1]) return;
    at<0, 3>(i, j, k);
  }
  template <int AC, int WD>
  HD void at(int i, int j, int k) const {
    const int I[3] = {i, j, k};
    if (wall_face(g, AC, I[AC])) return;
    const unsigned o = (unsigned)fidx(g, AC, i, j, k);
    const T old = ld(phi[AC], o), r = ld(rhs[AC], o);
    T diag;
    const T L = lphi<T, VT, DM, WD>(g, C, phi, AC, i, j, k, &diag);
    st(phi[AC], o, old + omega * (r - L) / diag);
  }
};

// face restriction of component a: 12/6-point (residuals, wall faces zeroed) or
// 4/2-point mean (coefficients), stag_mg_solver.f90:966-1170
template <typename T, int DM>
struct RestrictFn {
  using value_type = T;
  Geo gf, gc;
  const T* f;
  T* c;
  int a;
  Dec dc;
  int simple;
  int sf = 0, sc = 0;  // fine / coarse arrays in parity-split storage
  T* z = nullptr;      // when set, the coarse level's phi: zeroed at every face written
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) {
      const unsigned o = sc ? sidx(gc, ac, i, j, k) : (unsigned)t;
      st(c, o, eval(i, j, k, t, ac));
      if (z) st(z, o, T(0));
    });
  }
  HD void put(int i, int j, int k, long t, T v) const {
    const unsigned o = sc ? sidx(gc, a, i, j, k) : (unsigned)t;
    st(c, o, v);
    if (z) st(z, o, T(0));
  }
  HD T eval(int i, int j, int k, long, const int a) const {
    const int I[3] = {i, j, k};
    int tb[2];
    transverse(a, tb[0], tb[1]);
    const int ea0 = a == 0, ea1 = a == 1, ea2 = a == 2;
    const int fi = 2 * i, fj = 2 * j, fk = DM == 3 ? 2 * k : 0;
    T s0 = T(0), sp = T(0), sm = T(0);
#pragma unroll
    for (int q = 0; q < (1 << (DM - 1)); ++q) {
      int o[3] = {0, 0, 0};
      o[tb[0]] += q & 1;
      if (DM == 3) o[tb[1]] += (q >> 1) & 1;
      const int ii = fi + o[0], jj = fj + o[1], kk = fk + o[2];
      s0 += ld(f, lidx(gf, a, ii, jj, kk, sf));
      sp += vel_at(f, gf, a, ii + ea0, jj + ea1, kk + ea2, (const T*)nullptr, sf);  // unused when simple
      sm += vel_at(f, gf, a, ii - ea0, jj - ea1, kk - ea2, (const T*)nullptr, sf);
    }
    if (simple) return (DM == 3 ? T(0.25) : T(0.5)) * s0;
    if (wall_face(gc, a, I[a])) return T(0);
    if (DM == 3) return T(0.125) * s0 + T(0.0625) * sp + T(0.0625) * sm;
    return T(0.25) * s0 + T(0.125) * (sp + sm);
  }
};

// prolongation-add of component a (stag_mg_solver.f90:1365-1628), one fine face per item, natural
// storage (the one-workgroup coarse levels)
template <typename T, int DM>
struct ProlongFn {
  using value_type = T;
  Geo gf, gc;
  T* f;
  const T* c;
  int a;
  Dec dc;
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) { st(f, (unsigned)t, eval(i, j, k, t, ac)); });
  }
  // the new value of face (i,j,k), stored at offset t of f (wall faces unchanged); every
  // operand is loaded before any is used, and the even and odd sums are both formed, so the
  // loads are in flight together
  HD T eval(int i, int j, int k, long t, const int a) const {
    const int I[3] = {i, j, k};
    const T old = ld(f, (unsigned)t);
    int tb[2];
    transverse(a, tb[0], tb[1]);
    int C0[3] = {i / 2, j / 2, k / 2};
    int off[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) off[d] = (I[d] % 2 == 0) ? -1 : 1;
    auto P = [&](int na, int nb, int nc) {
      int J[3] = {C0[0], C0[1], C0[2]};
      J[a] += na;
      J[tb[0]] += nb;
      if (DM == 3) J[tb[1]] += nc;
      return vel_at(c, gc, a, J[0], J[1], J[2]);
    };
    const int ob = off[tb[0]], oc = DM == 3 ? off[tb[1]] : 0;
    const bool even = I[a] % 2 == 0;
    T add;
    if (DM == 3) {
      const T p000 = P(0, 0, 0), p0b0 = P(0, ob, 0), p00c = P(0, 0, oc), p0bc = P(0, ob, oc);
      const T p100 = P(1, 0, 0), p1b0 = P(1, ob, 0), p10c = P(1, 0, oc), p1bc = P(1, ob, oc);
      const T se = T(9.0 / 16.0) * p000 + T(3.0 / 16.0) * p0b0 + T(3.0 / 16.0) * p00c + T(1.0 / 16.0) * p0bc;
      const T so = T(9.0 / 32.0) * p000 + T(3.0 / 32.0) * p0b0 + T(3.0 / 32.0) * p00c + T(1.0 / 32.0) * p0bc +
                   T(9.0 / 32.0) * p100 + T(3.0 / 32.0) * p1b0 + T(3.0 / 32.0) * p10c + T(1.0 / 32.0) * p1bc;
      add = even ? se : so;
    } else {
      const T p000 = P(0, 0, 0), p0b0 = P(0, ob, 0), p100 = P(1, 0, 0), p1b0 = P(1, ob, 0);
      const T se = T(0.75) * p000 + T(0.25) * p0b0;
      const T so = T(0.375) * p000 + T(0.125) * p0b0 + T(0.375) * p100 + T(0.125) * p1b0;
      add = even ? se : so;
    }
    return wall_face(gf, a, I[a]) ? old : old + add;
  }
};

// Prolongation-add from the coarse side: one thread per coarse face (I,J,K) of component a adds
// to the 2^dm fine faces (2I, 2J+dj, 2K+dk) and, below the last coarse face along a, to the
// 2^dm faces (2I+1, ...) between I and I+1.  Those fine faces use the coarse values at I, I+1
// along a and J-1..J+1, K-1..K+1 across it, loaded once here (18 in 3-D) where ProlongFn loads
// 8 per fine face; each fine face gets ProlongFn's sums in ProlongFn's order.
template <typename T, int DM>
struct ProlongCFn {
  Geo gf, gc;
  T* f;
  const T* c;
  int a;
  Dec dcc;             // decoder of the coarse face array of component a
  int sf = 0, sc = 0;  // fine / coarse arrays in parity-split storage
  HD void operator()(long t) const {
    int I0, J0, K0;
    decode3(t, dcc, I0, J0, K0);
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) {
      int tb[2];
      transverse(ac, tb[0], tb[1]);
      const int C0[3] = {I0, J0, K0};
      constexpr int NK = DM == 3 ? 3 : 1;
      T P[2][3][NK];
#pragma unroll
      for (int na = 0; na < 2; ++na)
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
          for (int n = 0; n < NK; ++n) {
            int J[3] = {C0[0], C0[1], C0[2]};
            J[ac] += na;
            J[tb[0]] += m - 1;
            if (DM == 3) J[tb[1]] += n - 1;
            P[na][m][n] = vel_at(c, gc, ac, J[0], J[1], J[2], (const T*)nullptr, sc);
          }
      const int last = C0[ac] == fext(gc, ac, ac) - 1;  // no fine face beyond the last coarse one
#pragma unroll
      for (int fa = 0; fa < 2; ++fa) {
        if (fa == 1 && last) continue;
#pragma unroll
        for (int d0 = 0; d0 < 2; ++d0)
#pragma unroll
          for (int d1 = 0; d1 < (DM == 3 ? 2 : 1); ++d1) {
            int F[3] = {2 * C0[0], 2 * C0[1], 2 * C0[2]};
            F[ac] += fa;
            F[tb[0]] += d0;
            if (DM == 3) F[tb[1]] += d1;
            const unsigned o = lidx(gf, ac, F[0], F[1], F[2], sf);
            const T old = ld(f, o);
            const int m = d0 == 0 ? 0 : 2, n = DM == 3 ? (d1 == 0 ? 0 : 2) : 0, n1 = DM == 3 ? 1 : 0;
            T add;
            if (DM == 3) {
              const T p000 = P[0][1][n1], p0b0 = P[0][m][n1], p00c = P[0][1][n], p0bc = P[0][m][n];
              const T p100 = P[1][1][n1], p1b0 = P[1][m][n1], p10c = P[1][1][n], p1bc = P[1][m][n];
              const T se = T(9.0 / 16.0) * p000 + T(3.0 / 16.0) * p0b0 + T(3.0 / 16.0) * p00c + T(1.0 / 16.0) * p0bc;
              const T so = T(9.0 / 32.0) * p000 + T(3.0 / 32.0) * p0b0 + T(3.0 / 32.0) * p00c + T(1.0 / 32.0) * p0bc +
                           T(9.0 / 32.0) * p100 + T(3.0 / 32.0) * p1b0 + T(3.0 / 32.0) * p10c + T(1.0 / 32.0) * p1bc;
              add = fa == 0 ? se : so;
            } else {
              const T p000 = P[0][1][0], p0b0 = P[0][m][0], p100 = P[1][1][0], p1b0 = P[1][m][0];
              const T se = T(0.75) * p000 + T(0.25) * p0b0;
              const T so = T(0.375) * p000 + T(0.125) * p0b0 + T(0.375) * p100 + T(0.125) * p1b0;
              add = fa == 0 ? se : so;
            }
            st(f, o, wall_face(gf, ac, F[ac]) ? old : old + add);
          }
      }
    });
  }
};

// every component of a per-face functor F (with eval(i, j, k, offset, component) -> value and
// put(offset, value)) in one column launch
// over the nodal box of geometry g
template <class F, int DM>
struct AllFacesFn {
  Geo g;
  F f[3];
  FastDiv cx;
  // every component is evaluated (at the nearest face of its box where (i,j,k) lies outside
  // it, so that all loads stay in bounds) before any result is stored: the stores cannot then
  // hold back the next component's loads
  HD void operator()(int col, int k0, int k1) const {
    const int j = (int)fdiv(cx, (unsigned)col);
    const int i = col - j * (int)cx.d;
    for (int k = k0; k < k1; ++k) {
      typename F::value_type v[DM];
      int o[DM];
#pragma unroll
      for (int a = 0; a < DM; ++a) {
        const int ic = imin(i, fext(g, a, 0) - 1), jc = imin(j, fext(g, a, 1) - 1), kc = imin(k, fext(g, a, 2) - 1);
        o[a] = fidx(g, a, ic, jc, kc);
        v[a] = f[a].eval(ic, jc, kc, o[a], a);
      }
#pragma unroll
      for (int a = 0; a < DM; ++a)
        if (face_in(g, a, i, j, k)) f[a].put(i, j, k, o[a], v[a]);
    }
  }
};

// a per-face functor on every component and a per-cell functor, one column launch
template <class FF, class FC, int DM>
struct FacesCellFn {
  Geo g;
  FF f[3];
  FC c;
  FastDiv cx;
  HD void operator()(int col, int k0, int k1) const {
    const int j = (int)fdiv(cx, (unsigned)col);
    const int i = col - j * (int)cx.d;
    for (int k = k0; k < k1; ++k) {
#pragma unroll
      for (int a = 0; a < DM; ++a)
        if (face_in(g, a, i, j, k)) f[a].at(i, j, k, fidx(g, a, i, j, k));
      if (i < g.n[0] && j < g.n[1] && k < g.n[2]) c.at(i, j, k, cidx(g, i, j, k));
    }
  }
};

// multigrid entry: rhs = b (converted), phi = 0; with sp the level's arrays are in
// parity-split storage of geometry g, component a
template <typename T>
struct EnterFn {
  using value_type = T;
  const double* b;
  T* rhs;
  T* phi;
  Geo g{};
  int a = 0;
  int sp = 0;
  HD T eval(int, int, int, long t, int) const { return T(b[t]); }
  HD void put(int i, int j, int k, long t, T v) const {
    const unsigned m = sp ? sidx(g, a, i, j, k) : (unsigned)t;
    rhs[m] = v;
    phi[m] = T(0);
  }
};

// cell coefficient coarsening: mean of the 2^dm children (cc_restriction)
template <typename T, int DM>
struct CcRestrictFn {
  Geo gf, gc;
  const T* f;
  T* c;
  Dec dc;
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    T s = T(0);
#pragma unroll
    for (int q = 0; q < (1 << DM); ++q)
      s += f[cidx(gf, 2 * i + (q & 1), 2 * j + ((q >> 1) & 1), DM == 3 ? 2 * k + ((q >> 2) & 1) : 0)];
    c[t] = (DM == 3 ? T(0.125) : T(0.25)) * s;
  }
};

// edge coefficient coarsening (3-D: 2-point mean along the edge; 2-D nodes: injection)
template <typename T>
struct EdgeRestrictFn {
  Geo gf, gc;
  const T* f;
  T* c;
  int e;
  Dec dc;
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    if (gc.dm == 2) { c[t] = f[eidx(gf, e, 2 * i, 2 * j, 0)]; return; }
    const int i2 = 2 * i, j2 = 2 * j, k2 = 2 * k;
    if (e == 0) c[t] = T(0.5) * (f[eidx(gf, e, i2, j2, k2)] + f[eidx(gf, e, i2, j2, k2 + 1)]);
    else if (e == 1) c[t] = T(0.5) * (f[eidx(gf, e, i2, j2, k2)] + f[eidx(gf, e, i2, j2 + 1, k2)]);
    else c[t] = T(0.5) * (f[eidx(gf, e, i2, j2, k2)] + f[eidx(gf, e, i2 + 1, j2, k2)]);
  }
};

// ---------------------------------------------------------------------------------------
// Cell-centred MAC multigrid (this build's own design; the reference uses FBoxLib's
// ml_cc_solve, which is not in the reference tree).  Operator:
//   (L phi)_c = sum_d [ b_d(c+e_d)(phi(c+e_d)-phi(c)) - b_d(c)(phi(c)-phi(c-e_d)) ] / h^2
// with b = 1/alpha_fc on faces and b = 0 on wall faces (homogeneous Neumann).
template <typename T, int DM>
HD T mac_L(const Geo& g, T* const* b, T h2i, const T* phi, int i, int j, int k, T* diag) {
  const T pc = phi[cidx(g, i, j, k)];
  T s = T(0), dsum = T(0);
#pragma unroll
  for (int d = 0; d < DM; ++d) {
    const int e0 = d == 0, e1 = d == 1, e2 = d == 2;
    const T bp = b[d][fidx(g, d, i + e0, j + e1, k + e2)];
    const T bm = b[d][fidx(g, d, i, j, k)];
    s += bp * (cc_at(phi, g, i + e0, j + e1, k + e2) - pc) - bm * (pc - cc_at(phi, g, i - e0, j - e1, k - e2));
    dsum += bp + bm;
  }
  if (diag) *diag = -dsum * h2i;
  return s * h2i;
}

template <typename T, int DM>
struct MacGsFn {  // in place, one parity (consistent red-black colouring)
  Geo g;
  T* b[3];
  T h2i;
  T* phi;
  const T* rhs;
  int par;
  Dec dc;
  HD void operator()(long t) const {
    int i, j, k;
    if (!decode_color(t, dc, par, i, j, k)) return;
    const long o = cidx(g, i, j, k);
    const T old = phi[o], r = rhs[o];
    T diag;
    const T L = mac_L<T, DM>(g, b, h2i, phi, i, j, k, &diag);
    if (diag != T(0)) phi[o] = old + (r - L) / diag;
  }
};

template <typename T, int DM>
struct MacGsNewFn {  // phase 1 of the two-phase sweep: new values of one parity into tmp
  Geo g;
  T* b[3];
  T h2i;
  const T* phi;
  const T* rhs;
  T* tmp;
  int par;
  Dec dc;
  HD void operator()(long t) const {
    int i, j, k;
    if (!decode_color(t, dc, par, i, j, k)) return;
    T diag;
    const T L = mac_L<T, DM>(g, b, h2i, phi, i, j, k, &diag);
    const long o = cidx(g, i, j, k);
    tmp[o] = diag != T(0) ? phi[o] + (rhs[o] - L) / diag : phi[o];
  }
};

template <typename T>
struct CopyColorCellFn {  // phase 2: phi <- tmp on one parity
  Geo g;
  T* phi;
  const T* tmp;
  int par;
  Dec dc;
  HD void operator()(long t) const {
    int i, j, k;
    if (!decode_color(t, dc, par, i, j, k)) return;
    const long o = cidx(g, i, j, k);
    phi[o] = tmp[o];
  }
};

template <typename T, int DM>
struct MacResidRestrictFn {  // coarse rhs = mean of the children's residuals
  Geo gf, gc;
  T* b[3];
  T h2i;
  const T* phi;
  const T* rhs;
  T* rc;
  Dec dc;
  T* z = nullptr;  // when set, the coarse level's phi, zeroed
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    T s = T(0);
#pragma unroll
    for (int q = 0; q < (1 << DM); ++q) {
      const int ii = 2 * i + (q & 1), jj = 2 * j + ((q >> 1) & 1), kk = DM == 3 ? 2 * k + ((q >> 2) & 1) : 0;
      s += rhs[cidx(gf, ii, jj, kk)] - mac_L<T, DM>(gf, b, h2i, phi, ii, jj, kk, nullptr);
    }
    rc[t] = (DM == 3 ? T(0.125) : T(0.25)) * s;
    if (z) z[t] = T(0);
  }
};

template <typename T>
struct MacProlongFn {  // piecewise-constant prolongation-add
  Geo gf, gc;
  T* f;
  const T* c;
  Dec dc;
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    f[t] = f[t] + c[cidx(gc, i / 2, j / 2, k / 2)];
  }
};

// MAC face coefficients at level 0: 1/alpha, zero on wall faces
template <typename T>
struct MacCoef0Fn {
  Geo g;
  const double* alpha;
  T* b;
  int a;
  Dec dc;
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    const int I[3] = {i, j, k};
    b[t] = wall_face(g, a, I[a]) ? T(0) : T(1.0 / alpha[t]);
  }
};

// ---------------------------------------------------------------------------------------
// elementwise helpers
template <typename To, typename Ti>
struct ConvertFn {
  const Ti* in;
  To* out;
  Ti scale;
  HD void operator()(long t) const { out[t] = To(in[t] * scale); }
};
template <typename T>
struct FillFn {
  T* out;
  T v;
  HD void operator()(long t) const { out[t] = v; }
};
template <typename T>
struct ZeroWallFacesFn {  // multifab_physbc_domainvel on one component
  Geo g;
  T* u;
  int a;
  Dec dc;
  int sp = 0;  // u in parity-split storage
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    const int I[3] = {i, j, k};
    if (wall_face(g, a, I[a])) u[sp ? sidx(g, a, i, j, k) : (unsigned)t] = T(0);
  }
};

// copy with conversion between natural and parity-split storage of one face component:
// to_split: dst[split (i,j,k)] = To(src[t] * scale), else dst[t] = To(src[split (i,j,k)] * scale),
// t the natural offset of face (i,j,k)
template <typename Ti, typename To>
struct SplitCopyFn {
  Geo g;
  const Ti* src;
  To* dst;
  Ti scale;
  int a;
  Dec dc;
  int to_split;
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    const unsigned s = sidx(g, a, i, j, k);
    if (to_split) dst[s] = To(src[t] * scale);
    else dst[t] = To(src[s] * scale);
  }
};

// ---------------------------------------------------------------------------------------
// preconditioner glue (apply_precon.f90:103-195)
// mac_rhs = D x_u + b_p  (compute_div then multifab_plus_plus)
template <typename T, int DM>
struct DivPlusFn {
  Geo g;
  const T* u[3];
  const double* bp;  // may be null
  T* out;
  T dxinv;
  Dec dc;
  T* out2 = nullptr;  // optional second copy
  int sp = 0;         // u in parity-split storage
  T* z = nullptr;     // optional array zeroed (the MAC solve's initial phi)
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    T s = (u[0][lidx(g, 0, i + 1, j, k, sp)] - u[0][lidx(g, 0, i, j, k, sp)]) * dxinv +
          (u[1][lidx(g, 1, i, j + 1, k, sp)] - u[1][lidx(g, 1, i, j, k, sp)]) * dxinv;
    if (DM == 3) s += (u[2][lidx(g, 2, i, j, k + 1, sp)] - u[2][lidx(g, 2, i, j, k, sp)]) * dxinv;
    const T v = bp ? T(s + T(bp[t])) : s;
    out[t] = v;
    if (out2) out2[t] = v;
    if (z) z[t] = T(0);
  }
};

// x_u = phi0 - (1/alpha) G Phi (subtract_weighted_gradp, macproject.f90:318-362), FP64 output
template <typename T>
struct PreconFaceFn {
  Geo g;
  const T* phi0;     // multigrid solution, component a
  const T* Phi;      // MAC potential
  const double* alpha;
  double* x;
  double dxinv;
  int a;
  Dec dc;
  int sp = 0;  // phi0 in parity-split storage
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    at(i, j, k, t);
  }
  HD void at(int i, int j, int k, long t) const {
    const int I[3] = {i, j, k};
    if (wall_face(g, a, I[a])) { x[t] = 0.0; return; }
    const double gp = ((double)cc_at(Phi, g, i, j, k) -
                       (double)cc_at(Phi, g, i - (a == 0), j - (a == 1), k - (a == 2))) * dxinv;
    x[t] = (double)phi0[sp ? sidx(g, a, i, j, k) : (unsigned)t] - gp * (1.0 / alpha[t]);
  }
};

// multigrid rhs b_u - G x_p with compute_grad's wall factor 2 (x_p null: b_u alone)
template <typename T>
struct RhsGradFn {
  Geo g;
  const double* bu;
  const double* xp;
  T* out;
  double dxinv;
  int a;
  Dec dc;
  int sp = 0;  // out in parity-split storage
  HD void operator()(long t) const {
    int i = 0, j = 0, k = 0;
    if (xp || sp) decode3(t, dc, i, j, k);
    const unsigned m = sp ? sidx(g, a, i, j, k) : (unsigned)t;
    if (!xp) { out[m] = T(bu[t]); return; }
    const int I[3] = {i, j, k};
    const double fac = wall_face(g, a, I[a]) ? 2.0 * dxinv : dxinv;
    const double gp = (cc_at(xp, g, i, j, k) - cc_at(xp, g, i - (a == 0), j - (a == 1), k - (a == 2))) * fac;
    out[m] = T(bu[t] - gp);
  }
};

// L_alpha phi = D(alpha^-1 G phi) on every cell (the MAC operator of level 0; cc_applyop)
template <typename T, int DM>
struct MacApplyFn {
  Geo g;
  T* b[3];
  T h2i;
  const T* phi;
  T* out;
  Dec dc;
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    out[t] = mac_L<T, DM>(g, b, h2i, phi, i, j, k, nullptr);
  }
};

// precon_type +-4 pressure (apply_precon.f90:372-414): x_p = beta b_p - theta Phi, scaled by
// -1 / -2 / -4/3 (type 4) or 1 / 2 / 4/3 (type -4, where |visc_type| 1 keeps it unscaled), plus
// -gamma b_p / +gamma b_p for |visc_type| 3
template <typename T>
struct PreconCell4Fn {
  const double* bp;
  const T* Phi;
  const double* beta;
  const double* gamma;
  double* xp;
  double theta;
  int mac, avt, pt;
  HD void operator()(long t) const {
    const double tphi = mac ? (double)Phi[t] * theta : (double)Phi[t];
    double v = bp[t] * beta[t] - tphi;
    if (avt == 1 && pt == 4) v = v * -1.0;
    else if (avt == 2 && pt == 4) v = v * -2.0;
    else if (avt == 2 && pt == -4) v = v * 2.0;
    else if (avt == 3 && pt == 4) v = v * (-4.0 / 3.0) + (bp[t] * gamma[t]) * -1.0;
    else if (avt == 3 && pt == -4) v = v * (4.0 / 3.0) + bp[t] * gamma[t];
    xp[t] = v;
  }
};

// x_p = -chi mac_rhs + theta Phi, chi = beta (|vt| 1), 2 beta (|vt| 2), 4/3 beta + gamma (|vt| 3)
template <typename T>
struct PreconCellFn {
  const T* mac_rhs;
  const T* Phi;
  const double* beta;
  const double* gamma;
  double* xp;
  double theta;
  int avt;
  HD void at(int, int, int, long t) const { (*this)(t); }
  HD void operator()(long t) const {
    double v = -(double)mac_rhs[t];
    if (avt == 1) v = v * beta[t];
    else if (avt == 2) v = v * beta[t] * 2.0;
    else { const double tt = v * gamma[t]; v = v * beta[t] * (4.0 / 3.0) + tt; }
    xp[t] = v + (double)Phi[t] * theta;
  }
};

// apply_matrix faces: (theta alpha - A0) x_u + G x_p, zero on wall faces
template <int VT, int DM>
struct AmFaceFn {
  Geo g;
  Coefs<double> C;
  double* xu[3];
  const double* xp;
  double* out;
  double dxinv;
  int a;
  Dec dc;
  HD void operator()(long t) const {
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) {
      int i, j, k;
      decode3(t, dc, i, j, k);
      const int I[3] = {i, j, k};
      if (wall_face(g, ac, I[ac])) { out[t] = 0.0; return; }
      const double L = lphi<double, VT, DM>(g, C, xu, ac, i, j, k, nullptr);
      const double gp = (cc_at(xp, g, i, j, k) - cc_at(xp, g, i - (ac == 0), j - (ac == 1), k - (ac == 2))) * dxinv;
      out[t] = L + gp;
    });
  }
};

// apply_matrix faces with gmres_spatial_order = 4 (apply_matrix.f90:123-135) on a periodic
// box: stag_applyop_4o for visc_type 1 (stag_applyop.f90:1262-1510, the 4th-order Laplacian
// of each component: (theta alpha + K b/(12 h^2)) phi - (16 phi(i+1) + 16 phi(i-1) - phi(i+2)
// - phi(i-2) + ... over the dimensions) b/(12 h^2), K = 60 in 2-D and 90 in 3-D) plus
// compute_grad_4o (div_and_grad.f90:705-880: (-p(i+1) + 15 p(i) - 15 p(i-1) + p(i-2)) /
// (12 h), as written there), each term in the reference's order
template <int DM>
struct AmFace4Fn {
  Geo g;
  const double* alpha[3];
  const double* beta;  // constant (visc_type 1): the reference reads its first cell
  double theta;
  double* xu[3];
  const double* xp;
  double* out;
  double t12sq, t12;  // 1/(12 h^2), 1/(12 h)
  int a;
  Dec dc;
  HD void operator()(long t) const {
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) {
      int i, j, k;
      decode3(t, dc, i, j, k);
      const double* p = xu[ac];
      auto v = [&](int di, int dj, int dk) { return vel_at(p, g, ac, i + di, j + dj, k + dk); };
      double s = 16.0 * v(1, 0, 0);
      s = s + 16.0 * v(-1, 0, 0);
      s = s - 1.0 * v(2, 0, 0);
      s = s - 1.0 * v(-2, 0, 0);
      s = s + 16.0 * v(0, 1, 0);
      s = s + 16.0 * v(0, -1, 0);
      s = s - 1.0 * v(0, 2, 0);
      s = s - 1.0 * v(0, -2, 0);
      if (DM == 3) {
        s = s + 16.0 * v(0, 0, 1);
        s = s + 16.0 * v(0, 0, -1);
        s = s - 1.0 * v(0, 0, 2);
        s = s - 1.0 * v(0, 0, -2);
      }
      const double b = beta[0];
      const double K = DM == 3 ? 90.0 : 60.0;
      const double L = (alpha[ac][t] * theta + K * b * t12sq) * p[t] - s * b * t12sq;
      const int e0 = ac == 0, e1 = ac == 1, e2 = ac == 2;
      const double gp = (-cc_at(xp, g, i + e0, j + e1, k + e2) + 15.0 * cc_at(xp, g, i, j, k) -
                         15.0 * cc_at(xp, g, i - e0, j - e1, k - e2) + cc_at(xp, g, i - 2 * e0, j - 2 * e1, k - 2 * e2)) *
                        t12;
      out[t] = L + gp;
    });
  }
};

// compute_grad (div_and_grad.f90:417-703) with FOEXTRAP ghosts: zero on wall faces
struct GradFn {
  Geo g;
  const double* p;
  double* out;
  double dxinv;
  int a;
  Dec dc;
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    const int I[3] = {i, j, k};
    const double fac = wall_face(g, a, I[a]) ? 2.0 * dxinv : dxinv;
    out[t] = (cc_at(p, g, i, j, k) - cc_at(p, g, i - (a == 0), j - (a == 1), k - (a == 2))) * fac;
  }
};

// -D x_u (apply_matrix b_p); x_u wall faces are already zero
template <int DM>
struct NegDivFn {
  Geo g;
  const double* u[3];
  double* out;
  double dxinv;
  Dec dc;
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    double s = (u[0][fidx(g, 0, i + 1, j, k)] - u[0][fidx(g, 0, i, j, k)]) * dxinv +
               (u[1][fidx(g, 1, i, j + 1, k)] - u[1][fidx(g, 1, i, j, k)]) * dxinv;
    if (DM == 3) s += (u[2][fidx(g, 2, i, j, k + 1)] - u[2][fidx(g, 2, i, j, k)]) * dxinv;
    out[t] = -s;
  }
};

// ---------------------------------------------------------------------------------------
// On-device noise: a counter-based generator, so that every noise value is a pure function of
// (seed, array, position, stage) and kernels can regenerate it where they need it instead of
// reading a stored array.  Philox4x64-10 (Salmon, Moraes, Dror & Shaw, SC'11, "Parallel random
// numbers: as easy as 1, 2, 3"), the uniform ((w >> 11) + 1/2) 2^-53 in (0,1), and the normal
// quantile by Acklam's rational approximation refined with one Halley step on erfc (Acklam's
// own refinement), accurate to a few ulp.
struct Philox {
  unsigned long long k0, k1;
  HD static unsigned long long mulhi(unsigned long long a, unsigned long long b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (unsigned long long)(((unsigned __int128)a * b) >> 64);
#endif
  }
  HD void block(unsigned long long c[4]) const {
    unsigned long long ka = k0, kb = k1;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
      const unsigned long long lo0 = 0xD2E7470EE14C6C93ull * c[0], hi0 = mulhi(0xD2E7470EE14C6C93ull, c[0]);
      const unsigned long long lo1 = 0xCA5A826395121157ull * c[2], hi1 = mulhi(0xCA5A826395121157ull, c[2]);
      const unsigned long long n0 = hi1 ^ c[1] ^ ka, n2 = hi0 ^ c[3] ^ kb;
      c[0] = n0;
      c[1] = lo1;
      c[2] = n2;
      c[3] = lo0;
      ka += 0x9E3779B97F4A7C15ull;
      kb += 0xBB67AE8584CAA73Bull;
    }
  }
};

HD double normal_quantile(double p) {
  const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                       1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00};
  const double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                       6.680131188771972e+01, -1.328068155288572e+01};
  const double c[6] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                       -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00};
  const double d[4] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                       3.754408661907416e+00};
  double x;
  if (p < 0.02425) {
    const double q = sqrt(-2.0 * log(p));
    x = (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
        ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
  } else if (p > 0.97575) {
    const double q = sqrt(-2.0 * log(1.0 - p));
    x = -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
        ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
  } else {
    const double q = p - 0.5, r = q * q;
    x = (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
        (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.0);
  }
  // Halley step on Phi(x) - p, Phi(x) = erfc(-x / sqrt 2) / 2
  const double e = 0.5 * erfc(-x * 0.70710678118654752440) - p;
  const double u = e * 2.50662827463100050242 * exp(0.5 * x * x);
  return x - u / (1.0 + 0.5 * x * u);
}

// N(0,1) value `word` (0..3) of the Philox block at counter (pos, array, stage, 0)
struct NoiseGen {
  Philox ph;
  unsigned long long stage;
  HD double normal(long pos, int array, int word) const {
    unsigned long long c[4] = {(unsigned long long)pos, (unsigned long long)array, stage, 0ull};
    ph.block(c);
    const unsigned long long w = word == 0 ? c[0] : (word == 1 ? c[1] : (word == 2 ? c[2] : c[3]));
    const double u = ((double)(w >> 11) + 0.5) * 1.1102230246251565404e-16;
    return normal_quantile(u);
  }
};

// ---------------------------------------------------------------------------------------
// stochastic_m_fluxdiv (stochastic_m_fluxdiv.f90:42-602): f_a = sum_b d_b Sigma_ab with
// Sigma_cc = var W_cc sqrt(eta T), Sigma_ed = var W_ed sqrt(eta_ed T_ed), x sqrt(2) on no-slip
// wall edges and 0 on slip wall edges (mflux_bc), duplicated periodic edges take the value at
// index 0 (multifab_internal_sync), f = 0 on wall faces.
struct StochFn {
  Geo g;
  const double* Wcc;      // component a, cell
  const double* eta_cc;
  const double* T_cc;     // may be null
  const double* Wed[2];   // per transverse b (t = 0, 1): the edge noise array and component
  const double* eta_ed[2];
  const double* T_ed[2];
  int eid[2];
  double var, dxinv, sqrt2;
  double* out;
  int a;
  Dec dc;
  // on-device noise instead of Wcc / Wed: cell array 0 (word = component, x sqrt 2 in the
  // symmetric form), edge array 1 + e (word = ecomp[t])
  int gen = 0;
  NoiseGen ng;
  int ecomp[2] = {0, 0};
  double ccfac = 1.0;
  int incr = 0;  // add to out (the reference's `increment`) instead of overwriting it
  // Wcc / Wed hold the finished (scaled, synced, filtered) stresses: read them as they are
  int pre = 0;
  HD double w_cc(long o) const { return gen ? ccfac * ng.normal(o, 0, a) : Wcc[o]; }
  HD double w_ed(int t, long o) const { return gen ? ng.normal(o, 1 + eid[t], ecomp[t]) : Wed[t][o]; }
  HD double edge_flux(int t, int i, int j, int k) const {
    const int e = eid[t];
    int I[3] = {i, j, k};
    double fac = 1.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (!enodal(g.dm, e, d) || d >= g.dm) continue;
      if (I[d] == 0 || I[d] == g.n[d]) {
        const int code_wall = I[d] == 0 ? g.wlo[d] : g.whi[d];
        if (code_wall) {
          const int noslip = (I[d] == 0 ? g.slo[d] : g.shi[d]) < 0;
          fac *= noslip ? sqrt2 : 0.0;
        } else if (g.per[d] && I[d] == g.n[d]) {
          I[d] = 0;
        }
      }
    }
    const long o = eidx(g, e, I[0], I[1], I[2]);
    if (pre) return fac * Wed[t][o];
    const double tt = T_ed[t] ? T_ed[t][o] : 1.0;
    return fac * ((var * w_ed(t, o)) * sqrt(eta_ed[t][o] * tt));
  }
  HD double cc_flux(int i, int j, int k) const {
    int I[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int n = g.n[d];
      if (I[d] < 0) I[d] = g.per[d] ? I[d] + n : 0;
      else if (I[d] >= n) I[d] = g.per[d] ? I[d] - n : n - 1;
    }
    const long o = cidx(g, I[0], I[1], I[2]);
    if (pre) return Wcc[o];
    const double tt = T_cc ? T_cc[o] : 1.0;
    return (var * w_cc(o)) * sqrt(eta_cc[o] * tt);
  }
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    const int I[3] = {i, j, k};
    if (wall_face(g, a, I[a])) {
      if (!incr) out[t] = 0.0;
      return;
    }
    int tb[2];
    transverse(a, tb[0], tb[1]);
    // directions in x, y, z order (stoch_m_force_3d)
    double v = 0.0;
    for (int b = 0; b < g.dm; ++b) {
      if (b == a) {
        v += (cc_flux(i, j, k) - cc_flux(i - (a == 0), j - (a == 1), k - (a == 2))) * dxinv;
      } else {
        const int tt = (b == tb[0]) ? 0 : 1;
        v += (edge_flux(tt, i + (b == 0), j + (b == 1), k + (b == 2)) - edge_flux(tt, i, j, k)) * dxinv;
      }
    }
    out[t] = incr ? out[t] + v : v;
  }
};

// ---- inhomogeneous velocity boundary values -------------------------------------------------
// vel_bc_t of set_inhomogeneous_vel_bcs (multifab_physbc_stag.f90:2462-2549): the tangential
// values of component b on the walls normal to d, an array nodal in d and b (3-D: the edge
// array of that pair; 2-D: the node array), at slot d (2-D) or 2 d + (b > d ? b - 1 : b) (3-D)
HD int vbt_slot(int dm, int d, int b) { return dm == 2 ? d : 2 * d + (b > d ? b - 1 : b); }

// phi_bc of convert_to_homogeneous: zero except on the wall faces, which carry vel_bc_n
struct BcPhiFn {
  Geo g;
  const double* vbn;
  double* out;
  int a;
  Dec dc;
  HD void operator()(long t) const {
    int I[3];
    decode3(t, dc, I[0], I[1], I[2]);
    out[t] = wall_face(g, a, I[a]) ? vbn[t] : 0.0;
  }
};

// Component a of phi_bc seen by the stencil: stored values (wall faces, periodic images) or,
// beyond a wall in a transverse direction d, the ghost that the inhomogeneous fill gives a zero
// interior (physbc_macvel_{2d,3d}_inhomogeneous, :1118-1274, :1647-2013): 2 v at a no-slip
// wall, -dx v (lo) / +dx v (hi) at a slip wall, v read at the wall point of the same column.
// From a face that is not a wall face the stencil reads no corner ghost, so one offset applies.
struct BcGhostAcc {
  static constexpr int sp = 0;  // natural storage
  const double* p;
  const double* const* vbt;
  const Geo* g;
  int a;
  double dx;
  HD unsigned idx(int i, int j, int k) const { return (unsigned)fidx(*g, a, i, j, k); }
  HD double c(long oa) const { return p[oa]; }
  HD double n(int i, int j, int k) const {
    const Geo& G = *g;
    int I[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d == a || d >= G.dm || G.per[d]) continue;
      const bool lo = I[d] < 0, hi = I[d] >= G.n[d];
      if (!lo && !hi) continue;
      I[d] = lo ? 0 : G.n[d];
      const int e = G.dm == 2 ? 0 : edge_id(3, a, d);
#pragma unroll
      for (int q = 0; q < 3; ++q) I[q] = I[q] < 0 ? 0 : imin(I[q], eext(G, e, q) - 1);  // in bounds always
      const double v = vbt[vbt_slot(G.dm, d, a)][eidx(G, e, I[0], I[1], I[2])];
      const int sgn = lo ? G.slo[d] : G.shi[d];
      return sgn < 0 ? 2.0 * v : (lo ? -dx * v : dx * v);
    }
    return vel_at(p, G, a, i, j, k);
  }
};

// out_a += sgn (theta alpha - A0) phi_bc on the faces that are not wall faces (the lift of
// convert_to_homogeneous, convert_to_homogeneous.f90:61-90; stag_applyop leaves wall faces 0)
template <int VT, int DM>
struct BcLiftFn {
  Geo g;
  Coefs<double> C;
  double* phi[3];
  const double* vbt[6];
  double dx, sgn;
  double* out;
  int a;
  Dec dc;
  HD void operator()(long t) const {
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) {
      int i, j, k;
      decode3(t, dc, i, j, k);
      const int I[3] = {i, j, k};
      if (wall_face(g, ac, I[ac])) return;
      const BcGhostAcc A{phi[ac], vbt, &g, ac, dx};
      const double r = lphi_acc<double, VT, DM>(g, C, phi, ac, i, j, k, nullptr, A);
      out[t] += sgn * r;
    });
  }
};

// out += sgn D u on the cells
template <int DM>
struct DivAddFn {
  Geo g;
  const double* u[3];
  double* out;
  double dxinv, sgn;
  Dec dc;
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    double s = (u[0][fidx(g, 0, i + 1, j, k)] - u[0][fidx(g, 0, i, j, k)]) * dxinv +
               (u[1][fidx(g, 1, i, j + 1, k)] - u[1][fidx(g, 1, i, j, k)]) * dxinv;
    if (DM == 3) s += (u[2][fidx(g, 2, i, j, k + 1)] - u[2][fidx(g, 2, i, j, k)]) * dxinv;
    out[t] += sgn * s;
  }
};

// Cell fields of write_plotfile (src_lowMach/write_plotfile.f90:358-392) from a Stokes vector:
// component q of out is averaged_vel a (average_face_to_cc, convert_stag.f90:76-128:
// 0.5 (u(i) + u(i + e_a))), shifted_vel a (shift_face_to_cc, :914-950: u(i), the lower face),
// or the pressure, as listed in kind[q] = 0 / 1 / 2 with the component in comp[q]
struct PlotCcFn {
  Geo g;
  const double* u[3];
  const double* p;
  double* out;
  long nc;
  int nq;
  int kind[7], comp[7];
  Dec dc;
  HD void operator()(long t) const {
    int I[3];
    decode3(t, dc, I[0], I[1], I[2]);
    for (int q = 0; q < nq; ++q) {
      const int a = comp[q];
      double v;
      if (kind[q] == 2) {
        v = p[t];
      } else {
        const double lo = u[a][fidx(g, a, I[0], I[1], I[2])];
        if (kind[q] == 1) {
          v = lo;
        } else {
          const double hi = u[a][fidx(g, a, I[0] + (a == 0), I[1] + (a == 1), I[2] + (a == 2))];
          v = 0.5 * (lo + hi);
        }
      }
      out[q * nc + t] = v;
    }
  }
};

// The scaled stresses of stochastic_m_fluxdiv written out, for filtering: the cell stress of
// component F.a over the cells (mode < 0), or the edge stress that component F.a reads along
// transverse slot `mode` over the edges of that array (duplicated periodic edges take index 0,
// multifab_internal_sync)
struct StochStageFn {
  StochFn F;
  int mode;
  double* out;
  Dec dc;
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    out[t] = mode < 0 ? F.cc_flux(i, j, k) : F.edge_flux(mode, i, j, k);
  }
};

// One pass of filter_stoch_m_flux (stochastic_m_fluxdiv.f90:604-778) along axis d of a cell or
// edge array of extents dc.nx x dc.ny x (t / (nx ny)): a circular convolution of period n
// (= the cell count; index n of a nodal axis is index 0's duplicate and so reads the same
// neighbours). Requires n >= w. Taps and their order of summation are the reference's.
struct StochFilterFn {
  const double* in;
  double* out;
  long stride;  // element stride along d
  int d, n, w;
  Dec dc;
  HD void operator()(long t) const {
#pragma clang fp contract(off)
    int I[3];
    decode3(t, dc, I[0], I[1], I[2]);
    const int x = I[d];
    const long base = t - (long)x * stride;
    auto at = [&](int o) {
      int y = x + o;
      y = y < 0 ? y + n : (y >= n ? y - n : y);
      return in[base + (long)y * stride];
    };
    const double c = at(0);
    double v;
    if (w == 1) {
      v = c / 2 + at(-1) / 4 + at(1) / 4;
    } else if (w == 2) {
      v = 5 * c / 8 - at(-2) / 16 + at(-1) / 4 - at(2) / 16 + at(1) / 4;
    } else {
      v = 93 * c / 128 + 7 * at(-1) / 32 - 7 * at(-2) / 64 + at(-3) / 32 - at(-4) / 256 + 7 * at(1) / 32 -
          7 * at(2) / 64 + at(3) / 32 - at(4) / 256;
    }
    out[t] = v;
  }
};

// average_cc_to_face: cc given with one ghost layer ((n+2)^dm array); wall faces take the
// ghost value (convert_stag.f90:293-393)
struct AvgFaceFn {
  Geo g;
  const double* cc;  // ghosted
  double* out;
  int a;
  Dec dc;
  HD double G(int i, int j, int k) const {
    const int gx = g.n[0] + 2, gy = g.n[1] + 2;
    const int kk = g.dm == 3 ? k + 1 : 0;
    return cc[(i + 1) + (long)gx * ((j + 1) + (long)gy * kk)];
  }
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    const int I[3] = {i, j, k};
    const int mi = i - (a == 0), mj = j - (a == 1), mk = k - (a == 2);
    double v = 0.5 * (G(i, j, k) + G(mi, mj, mk));
    if (I[a] == 0 && g.wlo[a]) v = G(mi, mj, mk);
    if (I[a] == g.n[a] && g.whi[a]) v = G(i, j, k);
    out[t] = v;
  }
};

// average_cc_to_edge (3-D) / average_cc_to_node (2-D): 4-point mean; on a wall the mean of
// the two ghost cells beside the edge (later walls in direction order win)
struct AvgEdgeFn {
  Geo g;
  const double* cc;
  double* out;
  int e;
  Dec dc;
  HD double G(int i, int j, int k) const {
    const int gx = g.n[0] + 2, gy = g.n[1] + 2;
    const int kk = g.dm == 3 ? k + 1 : 0;
    return cc[(i + 1) + (long)gx * ((j + 1) + (long)gy * kk)];
  }
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    int p0, p1;
    if (g.dm == 2 || e == 0) { p0 = 0; p1 = 1; }
    else if (e == 1) { p0 = 0; p1 = 2; }
    else { p0 = 1; p1 = 2; }
    int I[3] = {i, j, k};
    int A[3] = {i, j, k}, B[3] = {i, j, k}, Cc[3] = {i, j, k};
    A[p0] -= 1; B[p1] -= 1; Cc[p0] -= 1; Cc[p1] -= 1;
    double v = 0.25 * (G(I[0], I[1], I[2]) + G(A[0], A[1], A[2]) + G(B[0], B[1], B[2]) + G(Cc[0], Cc[1], Cc[2]));
    for (int w = 0; w < 2; ++w) {
      const int d = w == 0 ? p0 : p1;
      const int o = w == 0 ? p1 : p0;
      for (int side = 0; side < 2; ++side) {
        const int wall = side == 0 ? g.wlo[d] : g.whi[d];
        if (!wall) continue;
        const int idx = side == 0 ? 0 : g.n[d];
        if (I[d] != idx) continue;
        int G0[3] = {i, j, k};
        G0[d] = side == 0 ? -1 : g.n[d];
        int G1[3] = {G0[0], G0[1], G0[2]};
        G1[o] -= 1;
        v = 0.5 * (G(G0[0], G0[1], G0[2]) + G(G1[0], G1[1], G1[2]));
      }
    }
    out[t] = v;
  }
};

// ---------------------------------------------------------------------------------------
// Stokes-vector inner-product weights: 1/2 on the two box-boundary faces of each component's
// nodal direction (norm_inner_product.f90:183-313), p_norm_weight^2 on cells (gmres.f90:358)
struct VecLayout {
  long off[5];  // start of face 0,1,2 and cell arrays, total length in off[4]
  int n[3];
  int dm;
  double pw2;
  // divisors n0 + 1, n0, n1 + 1, n0 n1 (every offset of a vector is below 2^31)
  FastDiv f_x1, f_x0, f_y1, f_xy;
  void init_div() {
    f_x1 = make_fastdiv((unsigned)(n[0] + 1));
    f_x0 = make_fastdiv((unsigned)n[0]);
    f_y1 = make_fastdiv((unsigned)(n[1] + 1));
    f_xy = make_fastdiv((unsigned)(n[0] * n[1]));
  }
  HD double weight(long e) const {
    if (e < off[1]) {
      const unsigned u = (unsigned)e;
      const int i = (int)(u - fdiv(f_x1, u) * f_x1.d);
      return (i == 0 || i == n[0]) ? 0.5 : 1.0;
    }
    if (e < off[2]) {
      const unsigned r = fdiv(f_x0, (unsigned)(e - off[1]));
      const int j = (int)(r - fdiv(f_y1, r) * f_y1.d);
      return (j == 0 || j == n[1]) ? 0.5 : 1.0;
    }
    if (e < off[3]) {
      const int k = (int)fdiv(f_xy, (unsigned)(e - off[2]));
      return (k == 0 || k == n[2]) ? 0.5 : 1.0;
    }
    return pw2;
  }
};


// ------------------------------------------------------------------------------------------
// Momentum time step for one species (src_lowMach/advance_timestep_inertial.f90,
// advance_timestep_overdamped.f90): the right-hand sides of its two Stokes solves, one face
// per thread, each term formed and added in the reference's order.

// sum over d of (f(c + e_d) - f(c)) / dx for component a at face (i,j,k): the flux of m_a = rho_a
// u_a through the d-faces of its control volume, f(c) = 1/4 (m_a(c - e_d) + m_a(c)) (u_d(c -
// e_a) + u_d(c)) (mk_advective_m_fluxdiv.f90:184-345; its update is minus this sum).  Momenta
// beyond the box are products of the wrapped / reflected density and velocity; at a wall they
// meet a zero normal velocity only.
template <int DM>
HD double adv_flux_sum(const Geo& g, const double* const* u, const double* const* rho, double dxinv, const int a,
                       int i, int j, int k) {
  const int ea0 = a == 0, ea1 = a == 1, ea2 = a == 2;
  auto m = [&](int ii, int jj, int kk) { return vel_at(rho[a], g, a, ii, jj, kk) * vel_at(u[a], g, a, ii, jj, kk); };
  double s = 0.0;
#pragma unroll
  for (int d = 0; d < DM; ++d) {
    const int e0 = d == 0, e1 = d == 1, e2 = d == 2;
    auto f = [&](int ci, int cj, int ck) {
      return 0.25 * (m(ci - e0, cj - e1, ck - e2) + m(ci, cj, ck)) *
             (vel_at(u[d], g, d, ci - ea0, cj - ea1, ck - ea2) + vel_at(u[d], g, d, ci, cj, ck));
    };
    const double t = (f(i + e0, j + e1, k + e2) - f(i, j, k)) * dxinv;
    s = d == 0 ? t : s + t;
  }
  return s;
}

// kind 0: inertial predictor (advance_timestep_inertial.f90:242-470), 1: inertial corrector
// (:640-810), 2: overdamped predictor or corrector (advance_timestep_overdamped.f90:200-335,
// 478-645).  The context's coefficients are those of the step's solves (inertial: alpha = rho,
// beta = eta/2, ...; overdamped: eta itself), and C.theta is 0, so L = lphi(v) is -A0 v with
// those coefficients.  In the inertial step the reference adds 1/2 A0(eta) v; halving is exact
// in binary floating point, so that term is -L bit for bit.
template <int VT, int DM>
struct StepRhsFn {
  Geo g;
  Coefs<double> C;
  double* v[3];          // velocity of the fluxes: v^n (predictor) or v* (corrector)
  const double* vn[3];   // v^n: mold = rho v^n
  const double* rho[3];  // face density (the context's alpha)
  const double* p;       // pressure: pi^n or pi*
  const double* f1[3];   // stochastic flux divergences (null when absent)
  const double* f2[3];
  double* adv[3];        // inertial: A^n, written by the predictor and read by the corrector
  double* dif[3];        // inertial: L(v^n), written by the predictor and read by the corrector
  double* out[3];
  double dxinv, dtinv;
  int kind, a;
  Dec dc;
  HD void operator()(long t) const {
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) {
      int i, j, k;
      decode3(t, dc, i, j, k);
      const int I[3] = {i, j, k};
      const bool wall = wall_face(g, ac, I[ac]);
      // compute_grad (div_and_grad.f90:417-703): 2/dx on a wall face, whose ghost is the cell
      const double gp = (cc_at(p, g, i, j, k) - cc_at(p, g, i - (ac == 0), j - (ac == 1), k - (ac == 2))) *
                        (wall ? 2.0 * dxinv : dxinv);
      // stag_applyop with alpha = 0; zero on wall faces like every stag_applyop result
      const double L = wall ? 0.0 : lphi<double, VT, DM>(g, C, v, ac, i, j, k, nullptr);
      double r;
      if (kind == 2) {
        r = -1.0 * gp + 0.0 * gp;         // multifab_saxpy_5(rhs, -1, gradpi, 0, gradpi)
        if (f1[0]) r = r + f1[ac][t];     // stochastic_m_fluxdiv, increment
        r = r - L;                        // diffusive_m_fluxdiv, increment
        out[ac][t] = r;
        return;
      }
      const double mold = rho[ac][t] * vn[ac][t];
      r = dtinv * mold + 0.0 * mold;      // multifab_saxpy_5(rhs, 1/dt, mold, 0, mold)
      r = r - gp;
      const double s = wall ? 0.0 : adv_flux_sum<DM>(g, v, rho, dxinv, ac, i, j, k);
      if (kind == 0) {
        const double A = wall ? 0.0 : -s;  // mk_advective_m_fluxdiv, new
        adv[ac][t] = A;
        r = r + A;
        dif[ac][t] = L;
        r = r + (-L);                     // + 1/2 A0^n v^n
        if (f1[0]) r = r + f1[ac][t];     // + div Sigma^n
        r = r + (-dtinv) * mold;          // - rho v^n / dt (mtemp, the same product)
        r = r + (-L);                     // + 1/2 A0^n v^n (mtemp)
      } else {
        const double A = wall ? 0.0 : adv[ac][t] - s;  // mk_advective_m_fluxdiv, increment
        r = r + 0.5 * A;
        r = r + (-dif[ac][t]);            // + 1/2 A0^n v^n
        if (f1[0]) {                      // + 1/2 (div Sigma^n + div Sigma^n')
          const double sc = f1[ac][t] + (f2[0] ? f2[ac][t] : f1[ac][t]);
          r = r + 0.5 * sc;
        }
        r = r + (-dtinv) * (rho[ac][t] * v[ac][t]);  // - rho v* / dt
        r = r + (-L);                     // + 1/2 A0^{n+1} v*
      }
      out[ac][t] = wall ? 0.0 : r;        // zero_edgeval_physical
    });
  }
};

// mk_advective_m_fluxdiv on its own: m_update = -sum (or m_update - sum), zero on wall faces
template <int DM>
struct AdvMFn {
  Geo g;
  const double* u[3];
  const double* rho[3];
  double* out;
  double dxinv;
  int a, increment;
  Dec dc;
  HD void operator()(long t) const {
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) {
      int i, j, k;
      decode3(t, dc, i, j, k);
      const int I[3] = {i, j, k};
      if (wall_face(g, ac, I[ac])) { out[t] = 0.0; return; }
      const double s = adv_flux_sum<DM>(g, u, rho, dxinv, ac, i, j, k);
      out[t] = increment ? out[t] - s : -s;
    });
  }
};

// inverse_diag_lap (src_gmres/inverse_diag_lap.f90:17-250): D = h^2 times the diagonal of A0
// on each face (Minv = 1/D, formed where the MAC coefficients are set up); visc_type -1 / -2
// / 1 / 2, each sum in the reference's order (2-D: the node above the face before the node at
// it; 3-D: the edge at the face before the one above)
template <int VT, int DM>
struct DiagSumFn {
  Geo g;
  const double* beta;
  const double* bed[3];
  double* out;
  int a;
  Dec dc;
  HD void operator()(long t) const {
    by_comp<DM>(a, [&](const int ac) __attribute__((always_inline)) {
      if (VT == 1 || VT == 2) {
        out[t] = (VT == 1 ? (DM == 3 ? 6.0 : 4.0) : (DM == 3 ? 8.0 : 6.0)) * beta[0];
        return;
      }
      int i, j, k;
      decode3(t, dc, i, j, k);
      const double bI = cc_at(beta, g, i, j, k);
      const double bM = cc_at(beta, g, i - (ac == 0), j - (ac == 1), k - (ac == 2));
      double s = VT == -2 ? 2.0 * bI + 2.0 * bM : bI + bM;
      int tb[2];
      transverse(ac, tb[0], tb[1]);
#pragma unroll
      for (int q = 0; q < DM - 1; ++q) {
        const int b = tb[q];
        const int e = edge_id(DM, ac, b);
        const double e0 = bed[e][eidx(g, e, i, j, k)];
        const double e1 = bed[e][eidx(g, e, i + (b == 0), j + (b == 1), k + (b == 2))];
        s = DM == 2 ? (s + e1) + e0 : (s + e0) + e1;
      }
      out[t] = s;
    });
  }
};

// visc_schur_complement pieces (apply_precon.f90:538-685): Minv (compute_grad of phi), the
// gradient with 2/h on a wall face (its ghost is the cell), then times Minv
template <typename T>
struct WGradFn {
  Geo g;
  const T* phi;
  const T* w;
  T* out;
  T dxinv;
  int a;
  Dec dc;
  HD void operator()(long t) const {
    int i, j, k;
    decode3(t, dc, i, j, k);
    const int I[3] = {i, j, k};
    const T fac = wall_face(g, a, I[a]) ? T(2) * dxinv : dxinv;
    out[t] = ((cc_at(phi, g, i, j, k) - cc_at(phi, g, i - (a == 0), j - (a == 1), k - (a == 2))) * fac) * w[t];
  }
};
template <typename T>
struct MulFn {
  T* x;
  const T* w;
  HD void operator()(long t) const { x[t] = x[t] * w[t]; }
};
// x_p = phi_mu (+ theta phi when the L_alpha solve ran)
template <typename T>
struct ViscOutFn {
  const T* pmu;
  const T* phi;
  double theta;
  int add;
  double* xp;
  HD void operator()(long t) const { xp[t] = add ? (double)pmu[t] + (double)phi[t] * theta : (double)pmu[t]; }
};

// x += dx over n elements (multifab_plus_plus_c on the velocity and the pressure)
struct AddFn {
  double* x;
  const double* d;
  HD void operator()(long t) const { x[t] = x[t] + d[t]; }
};

}  // namespace lmfhd
