// wadges_oracle.cpp -- CPU restatement of the WADGES right-hand-side / time-step
// hot path (volume flux differencing, surface ES flux + Lax-Friedrichs penalty,
// weight-adjusted inverse mass + LSRK update + entropy projection).
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is loaded by tests/,
// by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
// checker or as the reported CPU baseline, never as the product path.  The
// product is the HIP library under scratch_wadge_amd/csrc.
//
// What it restates (reference = ntan15/scratch_wadge, files relative to its
// root).  The reference's shipped kernels are a TESTING snapshot; the oracle
// follows the intended math ("O1"): the TESTING code paths of Euler3D.okl with
// only the log branch of logmean corrected, and the non-TESTING bodies of
// logmean/euler2d_flux in Euler2D.okl (the TESTING 2D flux is a constant stub).
//
//   logmean 3D   okl/Euler3D.okl:55-83   (series branch :63-66 as written,
//                                          log branch corrected to (aL+aR)f/ln(xi),
//                                          i.e. the intended :76-81)
//   logmean 2D   okl/Euler2D.okl:54-66
//   VU / UV 3D   okl/Euler3D.okl:24-51 (macros), :85-120 (TESTING VU :91-98)
//   VU / UV 2D   okl/Euler2D.okl:19-33, :69-98
//   flux 3D      okl/Euler3D.okl:122-162 (TESTING branch)
//   flux 2D      okl/Euler2D.okl:139-161
//   vol 3D       okl/Euler3D.okl:2123-2523 (euler_vol_3d_curved, TESTING 1/rho)
//   vol 2D       okl/Euler2D.okl:1112-1350 (euler_vol_2d_curved)
//   surf 3D      okl/Euler3D.okl:3242-3343 + lift :3390-3468 (euler_surf_3d)
//   surf 2D      okl/Euler2D.okl:1357-1495 (euler_surf_2d; LF there has no tau,
//                                           tau=1 reproduces it bit for bit)
//   update 3D    okl/Euler3D.okl:2527-2884 (euler_update_3d_curved, TESTING 1/J)
//   update 2D    okl/Euler2D.okl:1729-2282 (euler_update_2d_curved)
//   LSRK45       euler.c:2779-2796, stage loop euler.c:3691-3699
//
// Array layouts are the reference's (euler.c:2798-2877, operators_c.cpp:322-389):
//   Q, rhsQ, resQ  [k][fld][Nq]          Qf, rhsQf [k][fld][Nf]   (Nf = Nfq*Nfaces)
//   vgeo  [k][g][Nq]   vfgeo [k][g][Nf]   fgeo [k][g][Nf]   Jq [k][Nq]
//   mapPq [k][Nf] = field-strided index of the neighbour trace (euler.c:3158-3170)
//   Drq/Dsq/Dtq [i + j*Nq], VqLq [i + j*Nq], VfPq [i + j*Nf], VqPq [i + j*Nq]
//
// Every function is templated on the scalar type so that instantiating it with
// the FlopCount type below yields exact operation counts of the algorithm.
// Build with -ffp-contract=off so the arithmetic is evaluated exactly as written.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ----------------------------------------------------------------- flop counter

struct FlopTally {
  long long add = 0, mul = 0, div = 0, log = 0, exp = 0, pow = 0, sqrt = 0, cmp = 0;
  void reset() { *this = FlopTally(); }
  long long total() const { return add + mul + div + log + exp + pow + sqrt; }
};
FlopTally g_tally;

struct FlopCount {
  double v;
  FlopCount() : v(0) {}
  FlopCount(double x) : v(x) {}
};
inline FlopCount operator+(FlopCount a, FlopCount b) { g_tally.add++; return a.v + b.v; }
inline FlopCount operator-(FlopCount a, FlopCount b) { g_tally.add++; return a.v - b.v; }
inline FlopCount operator*(FlopCount a, FlopCount b) { g_tally.mul++; return a.v * b.v; }
inline FlopCount operator/(FlopCount a, FlopCount b) { g_tally.div++; return a.v / b.v; }
inline FlopCount operator-(FlopCount a) { return -a.v; }
inline FlopCount &operator+=(FlopCount &a, FlopCount b) { a = a + b; return a; }
inline bool operator<(FlopCount a, FlopCount b) { g_tally.cmp++; return a.v < b.v; }
inline bool operator>(FlopCount a, FlopCount b) { g_tally.cmp++; return a.v > b.v; }

inline double wlog(double x) { return std::log(x); }
inline double wexp(double x) { return std::exp(x); }
inline double wpow(double x, double y) { return std::pow(x, y); }
inline double wsqrt(double x) { return std::sqrt(x); }
inline double wabs(double x) { return std::fabs(x); }
inline double wmax(double a, double b) { return a > b ? a : b; }
inline FlopCount wlog(FlopCount x) { g_tally.log++; return std::log(x.v); }
inline FlopCount wexp(FlopCount x) { g_tally.exp++; return std::exp(x.v); }
inline FlopCount wpow(FlopCount x, FlopCount y) { g_tally.pow++; return std::pow(x.v, y.v); }
inline FlopCount wsqrt(FlopCount x) { g_tally.sqrt++; return std::sqrt(x.v); }
inline FlopCount wabs(FlopCount x) { return std::fabs(x.v); }
inline FlopCount wmax(FlopCount a, FlopCount b) { g_tally.cmp++; return a.v > b.v ? a : b; }

// ----------------------------------------------------------------- pointwise math

// 0: intended log mean.  1: the log branch as shipped in okl/Euler3D.okl:68,
// (aL+aR)/(ln(xi) f), which is off by 1/f^2.  Variant 1 exists only so the tests
// can plant that mistake in the oracle and show that their checks reject it.
int g_logmean_variant = 0;

// 0: the weight-adjusted update as written.  1: its first product without the
// projection, r = (-2 rhs) / J instead of (VqPq (-2 rhs)) / J -- a plausible slip
// that changes nothing on affine meshes (J constant per element).  Variant 1
// exists only so that the tests can plant that mistake in the oracle and show
// that the curved-mesh checks reject it.
int g_update_variant = 0;

// Log mean (aL - aR)/(ln aL - ln aR).
template <int DIM, typename T> inline T logmean(T aL, T aR) {
  if constexpr (DIM == 3) {
    // okl/Euler3D.okl:57-68 (TESTING) with the log branch written as intended.
    const T xi = aL / aR;
    const T f = (xi - T(1.0)) / (xi + T(1.0));
    const T u = f * f;
    const T u2 = u * u;
    if (wabs(u) < T(1.0e-4)) {
      const T F = T(105.0) + T(35.0) * u + T(21.0) * u2 + T(15.0) * u2 * u;
      return T(52.5) * (aL + aR) / F;
    }
    if (g_logmean_variant == 1) return (aL + aR) / (wlog(xi) * f);
    return (aL + aR) * f / wlog(xi);
  } else {
    // okl/Euler2D.okl:54-66.
    const T xi = aL / aR;
    const T f = (xi - T(1.0)) / (xi + T(1.0));
    const T u = f * f;
    const T u2 = u * u;
    T F;
    if (wabs(u) < T(1.0e-3)) {
      F = T(1.0) + u / T(3.0) + u2 / T(5.0) + u2 * u / T(7.0) + u2 * u2 / T(9.0);
    } else {
      F = wlog(xi) / T(2.0) / f;
    }
    return (aL + aR) / (T(2.0) * F);
  }
}

// Primitive state the flux needs: rho, velocity, E, beta = rho/(2p).
template <int DIM, typename T> struct Prim {
  T rho, vel[DIM], E, beta;
};

// Conservative -> primitive, the per-node loads of the volume kernels
// (Euler3D.okl:2252-2260 TESTING 1/rho multiply; Euler2D.okl:1127-1131 divide).
template <int DIM, typename T> inline Prim<DIM, T> prim_from(const T *U, double gamma) {
  Prim<DIM, T> p;
  p.rho = U[0];
  if constexpr (DIM == 3) {
    const T inv = T(1.0) / p.rho;
    for (int d = 0; d < 3; ++d) p.vel[d] = U[1 + d] * inv;
    p.E = U[4];
    const T uu = p.vel[0] * p.vel[0] + p.vel[1] * p.vel[1] + p.vel[2] * p.vel[2];
    const T pr = T(gamma - 1.0) * (p.E - T(0.5) * p.rho * uu);
    p.beta = p.rho / (T(2.0) * pr);
  } else {
    for (int d = 0; d < 2; ++d) p.vel[d] = U[1 + d] / p.rho;
    p.E = U[3];
    const T uu = p.vel[0] * p.vel[0] + p.vel[1] * p.vel[1];
    const T pr = T(gamma - 1.0) * (p.E - T(0.5) * p.rho * uu);
    p.beta = T(0.5) * p.rho / pr;
  }
  return p;
}

// Chandrashekar entropy-conservative two-point flux, F[d][fld].
template <int DIM, typename T>
inline void ec_flux(const Prim<DIM, T> &a, const Prim<DIM, T> &b, double gamma,
                    T F[DIM][DIM + 2]) {
  if constexpr (DIM == 3) {
    // okl/Euler3D.okl:128-162
    const T rholog = logmean<3>(a.rho, b.rho);
    const T betalog = logmean<3>(a.beta, b.beta);
    const T rhoavg = T(0.5) * (a.rho + b.rho);
    const T uavg = T(0.5) * (a.vel[0] + b.vel[0]);
    const T vavg = T(0.5) * (a.vel[1] + b.vel[1]);
    const T wavg = T(0.5) * (a.vel[2] + b.vel[2]);
    const T vnavg = T(2.0) * (uavg * uavg + vavg * vavg + wavg * wavg) -
                    (T(0.5) * (a.vel[0] * a.vel[0] + b.vel[0] * b.vel[0]) +
                     T(0.5) * (a.vel[1] * a.vel[1] + b.vel[1] * b.vel[1]) +
                     T(0.5) * (a.vel[2] * a.vel[2] + b.vel[2] * b.vel[2]));
    const T pa = rhoavg / (a.beta + b.beta);
    const T f4aux = rholog / (T(2.0 * (gamma - 1.0)) * betalog) + pa + T(0.5) * rholog * vnavg;
    F[0][0] = rholog * uavg;
    F[0][1] = F[0][0] * uavg + pa;
    F[0][2] = F[0][0] * vavg;
    F[0][3] = F[0][0] * wavg;
    F[0][4] = f4aux * uavg;
    F[1][0] = rholog * vavg;
    F[1][1] = F[1][0] * uavg;
    F[1][2] = F[1][0] * vavg + pa;
    F[1][3] = F[1][0] * wavg;
    F[1][4] = f4aux * vavg;
    F[2][0] = rholog * wavg;
    F[2][1] = F[2][0] * uavg;
    F[2][2] = F[2][0] * vavg;
    F[2][3] = F[2][0] * wavg + pa;
    F[2][4] = f4aux * wavg;
  } else {
    // okl/Euler2D.okl:139-161
    const T rholog = logmean<2>(a.rho, b.rho);
    const T rhoavg = T(0.5) * (a.rho + b.rho);
    const T uavg = T(0.5) * (a.vel[0] + b.vel[0]);
    const T vavg = T(0.5) * (a.vel[1] + b.vel[1]);
    const T vnavg = T(2.0) * (uavg * uavg + vavg * vavg) -
                    (T(0.5) * (a.vel[0] * a.vel[0] + b.vel[0] * b.vel[0]) +
                     T(0.5) * (a.vel[1] * a.vel[1] + b.vel[1] * b.vel[1]));
    const T beta_avg = T(0.5) * (a.beta + b.beta);
    const T pa = rhoavg / (T(2.0) * beta_avg);
    const T f4aux = rholog / (T(2.0 * (gamma - 1.0)) * logmean<2>(a.beta, b.beta)) + pa +
                    T(0.5) * rholog * vnavg;
    F[0][0] = rholog * uavg;
    F[1][0] = rholog * vavg;
    F[0][1] = F[0][0] * uavg + pa;
    F[1][1] = F[1][0] * uavg;
    F[0][2] = F[1][1];
    F[1][2] = F[1][0] * vavg + pa;
    F[0][3] = f4aux * uavg;
    F[1][3] = f4aux * vavg;
  }
}

// Conservative -> entropy variables (okl/Euler3D.okl:91-98 TESTING; Euler2D.okl:73-79).
template <int DIM, typename T> inline void VU(const T *U, T *V, double gamma) {
  const T g = T(gamma);
  const T rho = U[0];
  const T E = U[DIM + 1];
  T mm = U[1] * U[1] + U[2] * U[2];
  if constexpr (DIM == 3) mm = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
  const T rhoe = E - T(0.5) * mm / rho;
  const T s = wlog(T(gamma - 1.0) * rhoe / wpow(rho, g));
  if constexpr (DIM == 3) {
    const T invrhoe = rho / (E * rho - T(0.5) * mm);
    V[0] = ((-E * invrhoe) + (T(gamma + 1.0) - s));
    V[1] = U[1] * invrhoe;
    V[2] = U[2] * invrhoe;
    V[3] = U[3] * invrhoe;
    V[4] = (-rho) * invrhoe;
  } else {
    const T invrhoe = T(1.0) / rhoe;
    V[0] = (-E + rhoe * (T(gamma + 1.0) - s)) * invrhoe;
    V[1] = U[1] * invrhoe;
    V[2] = U[2] * invrhoe;
    V[3] = (-rho) * invrhoe;
  }
}

// Entropy -> conservative variables (okl/Euler3D.okl:110-120; Euler2D.okl:89-98).
template <int DIM, typename T> inline void UV(const T *V, T *U, double gamma) {
  const T g = T(gamma);
  const T vlast = V[DIM + 1];
  T vv = V[1] * V[1] + V[2] * V[2];
  if constexpr (DIM == 3) vv = V[1] * V[1] + V[2] * V[2] + V[3] * V[3];
  const T sV = g - V[0] + vv / (T(2.0) * vlast);
  const T rhoe = wpow(T(gamma - 1.0) / wpow(-vlast, g), T(1.0 / (gamma - 1.0))) *
                 wexp(-sV / T(gamma - 1.0));
  U[0] = rhoe * (-vlast);
  for (int d = 0; d < DIM; ++d) U[1 + d] = rhoe * (V[1 + d]);
  U[DIM + 1] = rhoe * (T(1.0) - vv / (T(2.0) * vlast));
}

// ----------------------------------------------------------------- kernels

struct Sizes {
  int Nq, Nfq, Nfaces, Nf;
};

// Averaged-metric contraction of one ordered volume-volume pair into divF
// (okl/Euler3D.okl:2317-2340; okl/Euler2D.okl:1160-1178).  gi/gj point at the
// DIM*DIM geofacs of the two nodes (g = r*DIM + d), D[r] = D_r(i,j).
template <int DIM, typename T>
inline void contract_vv(const T *gi, const T *gj, const T *D, const T F[DIM][DIM + 2],
                        T *divF) {
  constexpr int NF = DIM + 2;
  T Ga[DIM][DIM];
  for (int r = 0; r < DIM; ++r)
    for (int d = 0; d < DIM; ++d) Ga[r][d] = T(0.5) * (gi[r * DIM + d] + gj[r * DIM + d]);
  if constexpr (DIM == 3) {
    const T Dx = D[0] * Ga[0][0] + D[1] * Ga[1][0] + D[2] * Ga[2][0];
    const T Dy = D[0] * Ga[0][1] + D[1] * Ga[1][1] + D[2] * Ga[2][1];
    const T Dz = D[0] * Ga[0][2] + D[1] * Ga[1][2] + D[2] * Ga[2][2];
    for (int f = 0; f < NF; ++f) divF[f] += Dx * F[0][f] + Dy * F[1][f] + Dz * F[2][f];
  } else {
    for (int f = 0; f < NF; ++f)
      divF[f] += D[0] * (F[0][f] * Ga[0][0] + F[1][f] * Ga[0][1]) +
                 D[1] * (F[0][f] * Ga[1][0] + F[1][f] * Ga[1][1]);
  }
}

// Averaged-normal contraction of one volume-face pair for one coupling block
// (okl/Euler3D.okl:2383-2403 and :2488-2507; okl/Euler2D.okl:1208-1228, :1292-1312).
// w is the already scaled weight (+1/2 VqLq or -1/2 VfPq).
template <int DIM, typename T>
inline void contract_vf(const T *g1, const T *g2, const T *nref, T w, const T F[DIM][DIM + 2],
                        T *divF) {
  constexpr int NF = DIM + 2;
  T Ga[DIM][DIM];
  for (int r = 0; r < DIM; ++r)
    for (int d = 0; d < DIM; ++d) Ga[r][d] = T(0.5) * (g1[r * DIM + d] + g2[r * DIM + d]);
  if constexpr (DIM == 3) {
    const T nx = Ga[0][0] * nref[0] + Ga[1][0] * nref[1] + Ga[2][0] * nref[2];
    const T ny = Ga[0][1] * nref[0] + Ga[1][1] * nref[1] + Ga[2][1] * nref[2];
    const T nz = Ga[0][2] * nref[0] + Ga[1][2] * nref[1] + Ga[2][2] * nref[2];
    for (int f = 0; f < NF; ++f) divF[f] += w * (F[0][f] * nx + F[1][f] * ny + F[2][f] * nz);
  } else {
    for (int f = 0; f < NF; ++f)
      divF[f] += w * ((F[0][f] * Ga[0][0] + F[1][f] * Ga[0][1]) * nref[0] +
                      (F[0][f] * Ga[1][0] + F[1][f] * Ga[1][1]) * nref[1]);
  }
}

// One volume-face pair with the pair's work shared by its two coupling blocks:
// the geofac average, the averaged normal n = {G} nref and F.n are formed once
// and weighted into the volume row (w1 = +1/2 VqLq) and the face row
// (w2 = -1/2 VfPq).  Equal to contract_vf for both blocks up to rounding; this
// is the form whose operation count is the algorithmic work of a pair.
template <int DIM, typename T>
inline void contract_vf_pair(const T *gv, const T *gf, const T *nref, T w1, T w2,
                             const T F[DIM][DIM + 2], T *divV, T *divF) {
  constexpr int NF = DIM + 2;
  T Ga[DIM][DIM];
  for (int r = 0; r < DIM; ++r)
    for (int d = 0; d < DIM; ++d) Ga[r][d] = T(0.5) * (gv[r * DIM + d] + gf[r * DIM + d]);
  T n[DIM];
  for (int d = 0; d < DIM; ++d) {
    T v = Ga[0][d] * nref[0];
    for (int r = 1; r < DIM; ++r) v = v + Ga[r][d] * nref[r];
    n[d] = v;
  }
  for (int f = 0; f < NF; ++f) {
    T Fn = F[0][f] * n[0];
    for (int d = 1; d < DIM; ++d) Fn = Fn + F[d][f] * n[d];
    divV[f] += w1 * Fn;
    divF[f] += w2 * Fn;
  }
}

// Volume kernel for one element: rhsQ (vol rows) and rhsQf (face rows),
// okl/Euler3D.okl:2123-2523 / okl/Euler2D.okl:1112-1350.
template <int DIM, typename T>
void vol_element(const Sizes &S, double gamma, const T *vgeo, const T *vfgeo,
                 const T *nrstJ[DIM], const T *Drst[DIM], const T *VqLq, const T *VfPq,
                 const T *Uq, const T *Uf, T *rhsQ, T *rhsQf) {
  constexpr int NF = DIM + 2;
  constexpr int NG = DIM * DIM;
  const int Nq = S.Nq, Nf = S.Nf, Nfq = S.Nfq;
  std::vector<Prim<DIM, T>> pq(Nq), pf(Nf);
  std::vector<T> gq((size_t)Nq * NG), gf((size_t)Nf * NG);  // node-major geofacs
  for (int i = 0; i < Nq; ++i) {
    T U[NF];
    for (int f = 0; f < NF; ++f) U[f] = Uq[i + f * Nq];
    pq[i] = prim_from<DIM>(U, gamma);
    for (int g = 0; g < NG; ++g) gq[(size_t)i * NG + g] = vgeo[i + g * Nq];
  }
  for (int i = 0; i < Nf; ++i) {
    T U[NF];
    for (int f = 0; f < NF; ++f) U[f] = Uf[i + f * Nf];
    pf[i] = prim_from<DIM>(U, gamma);
    for (int g = 0; g < NG; ++g) gf[(size_t)i * NG + g] = vfgeo[i + g * Nf];
  }
  for (int i = 0; i < Nq; ++i) {
    T divF[NF];
    for (int f = 0; f < NF; ++f) divF[f] = T(0.0);
    for (int j = 0; j < Nq; ++j) {
      T F[DIM][NF];
      ec_flux<DIM>(pq[i], pq[j], gamma, F);
      T D[DIM];
      for (int r = 0; r < DIM; ++r) D[r] = Drst[r][i + j * Nq];
      contract_vv<DIM>(&gq[(size_t)i * NG], &gq[(size_t)j * NG], D, F, divF);
    }
    for (int j = 0; j < Nf; ++j) {
      const T w = T(0.5) * VqLq[i + j * Nq];
      T F[DIM][NF];
      ec_flux<DIM>(pq[i], pf[j], gamma, F);
      const int face = j / Nfq;
      T nref[DIM];
      for (int r = 0; r < DIM; ++r) nref[r] = nrstJ[r][face * Nfq];
      contract_vf<DIM>(&gq[(size_t)i * NG], &gf[(size_t)j * NG], nref, w, F, divF);
    }
    for (int f = 0; f < NF; ++f) rhsQ[i + f * Nq] = divF[f];
  }
  for (int i = 0; i < Nf; ++i) {
    T divF[NF];
    for (int f = 0; f < NF; ++f) divF[f] = T(0.0);
    const int face = i / Nfq;
    T nref[DIM];
    for (int r = 0; r < DIM; ++r) nref[r] = nrstJ[r][face * Nfq];
    for (int j = 0; j < Nq; ++j) {
      const T w = T(-0.5) * VfPq[i + j * Nf];
      T F[DIM][NF];
      ec_flux<DIM>(pf[i], pq[j], gamma, F);
      contract_vf<DIM>(&gf[(size_t)i * NG], &gq[(size_t)j * NG], nref, w, F, divF);
    }
    for (int f = 0; f < NF; ++f) rhsQf[i + f * Nf] = divF[f];
  }
}

// Surface kernel for one element (reads neighbour traces through mapPq).
template <int DIM, typename T>
void surf_element(const Sizes &S, double gamma, double tau, int k, const T *fgeo_k,
                  const uint32_t *mapPq_k, const T *VqLq, const T *Qf, const T *rhsQf_k, T *rhsQ_k) {
  constexpr int NF = DIM + 2;
  const int Nq = S.Nq, Nf = S.Nf;
  const T g = T(gamma);
  std::vector<T> sfS((size_t)NF * Nf);
  for (int i = 0; i < Nf; ++i) {
    const size_t idM = (size_t)i + (size_t)k * Nf * NF;
    const size_t idP = mapPq_k[i];
    T UM[NF], UP[NF];
    for (int f = 0; f < NF; ++f) {
      UM[f] = Qf[idM + (size_t)f * Nf];
      UP[f] = Qf[idP + (size_t)f * Nf];
    }
    T nJ[DIM];
    for (int d = 0; d < DIM; ++d) nJ[d] = fgeo_k[i + d * Nf];
    const T sJ = fgeo_k[i + DIM * Nf];
    Prim<DIM, T> pM, pP;
    T lamM, lamP;
    if constexpr (DIM == 3) {
      // okl/Euler3D.okl:3286-3323 (TESTING: multiply by 1/rho)
      const T invM = T(1.0) / UM[0], invP = T(1.0) / UP[0];
      pM.rho = UM[0];
      pP.rho = UP[0];
      for (int d = 0; d < 3; ++d) {
        pM.vel[d] = UM[1 + d] * invM;
        pP.vel[d] = UP[1 + d] * invP;
      }
      pM.E = UM[4];
      pP.E = UP[4];
      const T uuM = pM.vel[0] * pM.vel[0] + pM.vel[1] * pM.vel[1] + pM.vel[2] * pM.vel[2];
      const T uuP = pP.vel[0] * pP.vel[0] + pP.vel[1] * pP.vel[1] + pP.vel[2] * pP.vel[2];
      const T prM = T(gamma - 1.0) * (pM.E - T(0.5) * pM.rho * uuM);
      const T prP = T(gamma - 1.0) * (pP.E - T(0.5) * pP.rho * uuP);
      pM.beta = pM.rho / (T(2.0) * prM);
      pP.beta = pP.rho / (T(2.0) * prP);
      lamM = wsqrt(uuM) + wsqrt(g * prM / pM.rho);
      lamP = wsqrt(uuP) + wsqrt(g * prP / pP.rho);
    } else {
      // okl/Euler2D.okl:1402-1425
      pM = prim_from<2>(UM, gamma);
      pP = prim_from<2>(UP, gamma);
      const T uuM = pM.vel[0] * pM.vel[0] + pM.vel[1] * pM.vel[1];
      const T uuP = pP.vel[0] * pP.vel[0] + pP.vel[1] * pP.vel[1];
      const T prM = T(gamma - 1.0) * (pM.E - T(0.5) * pM.rho * uuM);
      const T prP = T(gamma - 1.0) * (pP.E - T(0.5) * pP.rho * uuP);
      lamM = wsqrt(uuM) + wsqrt(g * prM / pM.rho);
      lamP = wsqrt(uuP) + wsqrt(g * prP / pP.rho);
    }
    T F[DIM][NF];
    ec_flux<DIM>(pM, pP, gamma, F);
    const T Lfc = T(tau * 0.25) * wmax(lamM, lamP) * sJ;
    for (int f = 0; f < NF; ++f) {
      T fn;
      if constexpr (DIM == 3)
        fn = F[0][f] * nJ[0] + F[1][f] * nJ[1] + F[2][f] * nJ[2];
      else
        fn = F[0][f] * nJ[0] + F[1][f] * nJ[1];
      sfS[(size_t)f * Nf + i] = T(0.5) * fn + rhsQf_k[i + f * Nf] - Lfc * (UP[f] - UM[f]);
    }
  }
  for (int i = 0; i < Nq; ++i) {
    T val[NF];
    for (int f = 0; f < NF; ++f) val[f] = T(0.0);
    for (int j = 0; j < Nf; ++j) {
      const T L = VqLq[i + j * Nq];
      for (int f = 0; f < NF; ++f) val[f] += L * sfS[(size_t)f * Nf + j];
    }
    for (int f = 0; f < NF; ++f) rhsQ_k[i + f * Nq] += val[f];
  }
}

// Update kernel for one element (WADG inverse mass, LSRK, entropy projection).
template <int DIM, typename T>
void update_element(const Sizes &S, double gamma, const T *Jq_k, const T *VqPq, const T *VfPq,
                    T fa, T fb, T fdt, T *rhsQ_k, T *resQ_k, T *Q_k, T *Qf_k) {
  constexpr int NF = DIM + 2;
  const int Nq = S.Nq, Nf = S.Nf;
  std::vector<T> a((size_t)NF * Nq), b((size_t)NF * Nq);
  auto matvec = [&](const T *in, T *out, T scale) {
    for (int i = 0; i < Nq; ++i) {
      T r[NF];
      for (int f = 0; f < NF; ++f) r[f] = T(0.0);
      for (int j = 0; j < Nq; ++j) {
        const T P = scale * VqPq[i + j * Nq];
        for (int f = 0; f < NF; ++f) r[f] += P * in[j + f * Nq];
      }
      for (int f = 0; f < NF; ++f) out[i + f * Nq] = r[f];
    }
  };
  // r = VqPq (-2 rhs) / J   (okl/Euler3D.okl:2586-2611, TESTING: times 1/J)
  for (int i = 0; i < Nq; ++i) {
    T r[NF];
    for (int f = 0; f < NF; ++f) r[f] = T(0.0);
    for (int j = 0; j < Nq; ++j) {
      const T P = T(-2.0) * VqPq[i + j * Nq];
      for (int f = 0; f < NF; ++f) r[f] += P * rhsQ_k[j + f * Nq];
    }
    if (g_update_variant == 1)
      for (int f = 0; f < NF; ++f) r[f] = T(-2.0) * rhsQ_k[i + f * Nq];
    if constexpr (DIM == 3) {
      const T invJ = T(1.0) / Jq_k[i];
      for (int f = 0; f < NF; ++f) a[i + f * Nq] = r[f] * invJ;
    } else {
      for (int f = 0; f < NF; ++f) a[i + f * Nq] = r[f] / Jq_k[i];
    }
  }
  // r = VqPq a; res = fa res + dt r; Q += fb res; w = v(Q) J   (:2621-2703)
  for (int i = 0; i < Nq; ++i) {
    T r[NF];
    for (int f = 0; f < NF; ++f) r[f] = T(0.0);
    for (int j = 0; j < Nq; ++j) {
      const T P = VqPq[i + j * Nq];
      for (int f = 0; f < NF; ++f) r[f] += P * a[j + f * Nq];
    }
    T U[NF], V[NF];
    for (int f = 0; f < NF; ++f) {
      const T res = fa * resQ_k[i + f * Nq] + fdt * r[f];
      resQ_k[i + f * Nq] = res;
      U[f] = Q_k[i + f * Nq] + fb * res;
      Q_k[i + f * Nq] = U[f];
    }
    VU<DIM>(U, V, gamma);
    for (int f = 0; f < NF; ++f) b[i + f * Nq] = V[f] * Jq_k[i];
  }
  // a = VqPq b / J ; b = VqPq a    (:2706-2783)
  for (int i = 0; i < Nq; ++i) {
    T r[NF];
    for (int f = 0; f < NF; ++f) r[f] = T(0.0);
    for (int j = 0; j < Nq; ++j) {
      const T P = VqPq[i + j * Nq];
      for (int f = 0; f < NF; ++f) r[f] += P * b[j + f * Nq];
    }
    if constexpr (DIM == 3) {
      const T invJ = T(1.0) / Jq_k[i];
      for (int f = 0; f < NF; ++f) a[i + f * Nq] = r[f] * invJ;
    } else {
      for (int f = 0; f < NF; ++f) a[i + f * Nq] = r[f] / Jq_k[i];
    }
  }
  matvec(a.data(), b.data(), T(1.0));
  // Qf = u(VfPq vt), rhsQ = u(VqPq vt)   (:2788-2880)
  for (int i = 0; i < Nf; ++i) {
    T V[NF], U[NF];
    for (int f = 0; f < NF; ++f) V[f] = T(0.0);
    for (int j = 0; j < Nq; ++j) {
      const T P = VfPq[i + j * Nf];
      for (int f = 0; f < NF; ++f) V[f] += P * b[j + f * Nq];
    }
    UV<DIM>(V, U, gamma);
    for (int f = 0; f < NF; ++f) Qf_k[i + f * Nf] = U[f];
  }
  for (int i = 0; i < Nq; ++i) {
    T V[NF], U[NF];
    for (int f = 0; f < NF; ++f) V[f] = T(0.0);
    for (int j = 0; j < Nq; ++j) {
      const T P = VqPq[i + j * Nq];
      for (int f = 0; f < NF; ++f) V[f] += P * b[j + f * Nq];
    }
    UV<DIM>(V, U, gamma);
    for (int f = 0; f < NF; ++f) rhsQ_k[i + f * Nq] = U[f];
  }
}

Sizes make_sizes(int dim, int Nq, int Nfq) {
  Sizes S;
  S.Nq = Nq;
  S.Nfq = Nfq;
  S.Nfaces = dim + 1;
  S.Nf = Nfq * S.Nfaces;
  return S;
}

int g_threads = 1;

}  // namespace

// ===================================================================== C ABI

extern "C" {

typedef struct {
  int dim;       // 2 (triangles) or 3 (tetrahedra)
  int Nq;        // volume quadrature points per element
  int Nfq;       // quadrature points per face
  double gamma;  // ratio of specific heats
  double tau;    // Lax-Friedrichs penalty scale
} wo_params;

void wo_set_threads(int n) { g_threads = n > 0 ? n : 1; }

// Test hook: 0 = intended 3D log mean, 1 = the as-shipped log branch (see logmean).
void wo_set_logmean_variant(int v) { g_logmean_variant = v; }
void wo_set_update_variant(int v) { g_update_variant = v; }

int wo_get_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

double wo_logmean(int dim, double aL, double aR) {
  return dim == 3 ? logmean<3>(aL, aR) : logmean<2>(aL, aR);
}

// F is [dim][dim+2]; UL/UR are conservative states.
void wo_ec_flux(int dim, double gamma, const double *UL, const double *UR, double *F) {
  if (dim == 3) {
    auto a = prim_from<3>(UL, gamma), b = prim_from<3>(UR, gamma);
    double Fl[3][5];
    ec_flux<3>(a, b, gamma, Fl);
    std::memcpy(F, Fl, sizeof(Fl));
  } else {
    auto a = prim_from<2>(UL, gamma), b = prim_from<2>(UR, gamma);
    double Fl[2][4];
    ec_flux<2>(a, b, gamma, Fl);
    std::memcpy(F, Fl, sizeof(Fl));
  }
}

void wo_VU(int dim, double gamma, const double *U, double *V) {
  if (dim == 3) VU<3>(U, V, gamma); else VU<2>(U, V, gamma);
}

void wo_UV(int dim, double gamma, const double *V, double *U) {
  if (dim == 3) UV<3>(V, U, gamma); else UV<2>(V, U, gamma);
}

// Volume kernel over K elements; reads the projected state from rhsQ and
// overwrites it with the volume RHS (same in-place contract as the reference).
int wo_vol(const wo_params *p, int K, const double *vgeo, const double *vfgeo,
           const double *nrJ, const double *nsJ, const double *ntJ, const double *Drq,
           const double *Dsq, const double *Dtq, const double *VqLq, const double *VfPq,
           const double *Qf, double *rhsQ, double *rhsQf) {
  const Sizes S = make_sizes(p->dim, p->Nq, p->Nfq);
  const int dim = p->dim;
  const int NF = dim + 2, NG = dim * dim;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int k = 0; k < K; ++k) {
    std::vector<double> Uq(rhsQ + (size_t)k * S.Nq * NF, rhsQ + (size_t)(k + 1) * S.Nq * NF);
    const double *Uf = Qf + (size_t)k * S.Nf * NF;
    if (dim == 3) {
      const double *n[3] = {nrJ, nsJ, ntJ};
      const double *D[3] = {Drq, Dsq, Dtq};
      vol_element<3, double>(S, p->gamma, vgeo + (size_t)k * S.Nq * NG, vfgeo + (size_t)k * S.Nf * NG,
                             n, D, VqLq, VfPq, Uq.data(), Uf, rhsQ + (size_t)k * S.Nq * NF,
                             rhsQf + (size_t)k * S.Nf * NF);
    } else {
      const double *n[2] = {nrJ, nsJ};
      const double *D[2] = {Drq, Dsq};
      vol_element<2, double>(S, p->gamma, vgeo + (size_t)k * S.Nq * NG, vfgeo + (size_t)k * S.Nf * NG,
                             n, D, VqLq, VfPq, Uq.data(), Uf, rhsQ + (size_t)k * S.Nq * NF,
                             rhsQf + (size_t)k * S.Nf * NF);
    }
  }
  return 0;
}

int wo_surf(const wo_params *p, int K, const double *fgeo, const uint32_t *mapPq,
            const double *VqLq, const double *Qf, const double *rhsQf, double *rhsQ) {
  const Sizes S = make_sizes(p->dim, p->Nq, p->Nfq);
  const int dim = p->dim;
  const int NF = dim + 2, NFG = dim + 1;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int k = 0; k < K; ++k) {
    if (dim == 3)
      surf_element<3, double>(S, p->gamma, p->tau, k, fgeo + (size_t)k * S.Nf * NFG,
                              mapPq + (size_t)k * S.Nf, VqLq, Qf, rhsQf + (size_t)k * S.Nf * NF,
                              rhsQ + (size_t)k * S.Nq * NF);
    else
      surf_element<2, double>(S, p->gamma, p->tau, k, fgeo + (size_t)k * S.Nf * NFG,
                              mapPq + (size_t)k * S.Nf, VqLq, Qf, rhsQf + (size_t)k * S.Nf * NF,
                              rhsQ + (size_t)k * S.Nq * NF);
  }
  return 0;
}

int wo_update(const wo_params *p, int K, const double *Jq, const double *VqPq, const double *VfPq,
              double fa, double fb, double fdt, double *rhsQ, double *resQ, double *Q, double *Qf) {
  const Sizes S = make_sizes(p->dim, p->Nq, p->Nfq);
  const int dim = p->dim;
  const int NF = dim + 2;
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int k = 0; k < K; ++k) {
    const size_t oq = (size_t)k * S.Nq * NF, of = (size_t)k * S.Nf * NF;
    if (dim == 3)
      update_element<3, double>(S, p->gamma, Jq + (size_t)k * S.Nq, VqPq, VfPq, fa, fb, fdt,
                                rhsQ + oq, resQ + oq, Q + oq, Qf + of);
    else
      update_element<2, double>(S, p->gamma, Jq + (size_t)k * S.Nq, VqPq, VfPq, fa, fb, fdt,
                                rhsQ + oq, resQ + oq, Q + oq, Qf + of);
  }
  return 0;
}

// Exact operation counts (add, mul, div, log, exp, pow, sqrt) of the pieces the
// algorithmic flop model is built from:
//   out[0] c_F   : one two-point flux evaluation from primitive states
//   out[1] c_node: one conservative->primitive conversion (per volume/face node)
//   out[2] c_vv  : contraction work of one ordered volume-volume pair
//   out[3] c_vf  : contraction work of one volume-face pair, both coupling blocks,
//                  the average, normal and F.n shared by the blocks (contract_vf_pair)
//   out[4] vol   : whole reference-structured volume kernel for one element
//   out[5] surf  : whole surface kernel for one element
//   out[6] update: whole update kernel for one element
// plus transcendental splits out[7..] = {log, exp, pow, sqrt, div} of c_F, and
// out[12] = max |contract_vf_pair - contract_vf of both blocks| on that pair.
void wo_flop_counts(int dim, int Nq, int Nfq, double *out) {
  const Sizes S = make_sizes(dim, Nq, Nfq);
  const double gamma = 1.4;
  auto run = [&](auto DIMc) {
    constexpr int DIM = decltype(DIMc)::value;
    constexpr int NF = DIM + 2;
    using T = FlopCount;
    Prim<DIM, T> a, b;
    a.rho = 1.0; b.rho = 1.3;
    for (int d = 0; d < DIM; ++d) { a.vel[d] = 0.1 * (d + 1); b.vel[d] = -0.2 * (d + 1); }
    a.E = 3.0; b.E = 3.5; a.beta = 0.4; b.beta = 0.45;  // log branches of both log means
    T F[DIM][NF];
    g_tally.reset();
    ec_flux<DIM>(a, b, gamma, F);
    out[0] = (double)g_tally.total();
    out[7] = (double)g_tally.log; out[8] = (double)g_tally.exp; out[9] = (double)g_tally.pow;
    out[10] = (double)g_tally.sqrt; out[11] = (double)g_tally.div;
    T U[NF];
    U[0] = 1.0;
    for (int d = 0; d < DIM; ++d) U[1 + d] = 0.1;
    U[DIM + 1] = 3.0;
    g_tally.reset();
    (void)prim_from<DIM>(U, gamma);
    out[1] = (double)g_tally.total();

    // contraction of one ordered vol-vol pair, and one vol-face pair in both blocks
    {
      const int NG = DIM * DIM;
      T g1[NG], g2[NG], D[DIM], nref[DIM], divF[NF];
      for (int g = 0; g < NG; ++g) { g1[g] = 0.3 + g; g2[g] = 0.2 * g; }
      for (int r = 0; r < DIM; ++r) { D[r] = 0.1 * (r + 1); nref[r] = 1.0; }
      for (int f = 0; f < NF; ++f) divF[f] = 0.0;
      g_tally.reset();
      contract_vv<DIM>(g1, g2, D, F, divF);
      out[2] = (double)g_tally.total();
      // both blocks of one pair, the pair work counted once (contract_vf_pair);
      // out[12]: its largest deviation from the two per-block contractions
      const T w1 = T(0.5) * T(0.2), w2 = T(-0.5) * T(0.3);
      T dV[NF], dF[NF], eV[NF], eF[NF];
      for (int f = 0; f < NF; ++f) dV[f] = dF[f] = eV[f] = eF[f] = 0.0;
      g_tally.reset();
      contract_vf_pair<DIM>(g1, g2, nref, w1, w2, F, dV, dF);
      out[3] = (double)g_tally.total();
      contract_vf<DIM>(g1, g2, nref, w1, F, eV);
      contract_vf<DIM>(g2, g1, nref, w2, F, eF);
      double dev = 0.0;
      for (int f = 0; f < NF; ++f)
        dev = std::max(dev, std::max(std::fabs(dV[f].v - eV[f].v), std::fabs(dF[f].v - eF[f].v)));
      out[12] = dev;
    }
    // one element of each kernel with synthetic data (values only steer branches)
    const int NG = DIM * DIM;
    std::vector<T> vgeo((size_t)S.Nq * NG, T(0.3)), vfgeo((size_t)S.Nf * NG, T(0.2)),
        nr(S.Nf, T(1.0)), ns(S.Nf, T(0.5)), nt(S.Nf, T(-1.0)), D((size_t)S.Nq * S.Nq, T(0.1)),
        VqLq((size_t)S.Nq * S.Nf, T(0.05)), VfPq((size_t)S.Nf * S.Nq, T(0.05)),
        VqPq((size_t)S.Nq * S.Nq, T(0.02)), Uq((size_t)S.Nq * NF), Uf((size_t)S.Nf * NF),
        rq((size_t)S.Nq * NF), rf((size_t)S.Nf * NF), Jq(S.Nq, T(1.5)),
        fgeo((size_t)S.Nf * (DIM + 1), T(0.7));
    for (int i = 0; i < S.Nq; ++i) {
      Uq[i] = 1.0 + 0.1 * i;
      for (int d = 0; d < DIM; ++d) Uq[i + (1 + d) * S.Nq] = 0.05 * d;
      Uq[i + (DIM + 1) * S.Nq] = 3.0 + 0.2 * i;
    }
    for (int i = 0; i < S.Nf; ++i) {
      Uf[i] = 1.05 + 0.07 * i;
      for (int d = 0; d < DIM; ++d) Uf[i + (1 + d) * S.Nf] = 0.02 * d;
      Uf[i + (DIM + 1) * S.Nf] = 3.1 + 0.15 * i;
    }
    const T *n[DIM];
    const T *Dm[DIM];
    n[0] = nr.data();
    n[1] = ns.data();
    if constexpr (DIM == 3) n[2] = nt.data();
    for (int d = 0; d < DIM; ++d) Dm[d] = D.data();
    g_tally.reset();
    vol_element<DIM, T>(S, gamma, vgeo.data(), vfgeo.data(), n, Dm, VqLq.data(), VfPq.data(),
                        Uq.data(), Uf.data(), rq.data(), rf.data());
    out[4] = (double)g_tally.total();
    std::vector<uint32_t> map(S.Nf);
    for (int i = 0; i < S.Nf; ++i) map[i] = (uint32_t)i;
    g_tally.reset();
    surf_element<DIM, T>(S, gamma, 1.0, 0, fgeo.data(), map.data(), VqLq.data(), Uf.data(),
                         rf.data(), rq.data());
    out[5] = (double)g_tally.total();
    std::vector<T> res((size_t)S.Nq * NF, T(0.0)), Qv(Uq), Qfv((size_t)S.Nf * NF);
    for (size_t i = 0; i < rq.size(); ++i) rq[i] = T(0.01);
    g_tally.reset();
    update_element<DIM, T>(S, gamma, Jq.data(), VqPq.data(), VfPq.data(), T(0.5), T(0.3), T(0.01),
                           rq.data(), res.data(), Qv.data(), Qfv.data());
    out[6] = (double)g_tally.total();
  };
  if (dim == 3) run(std::integral_constant<int, 3>()); else run(std::integral_constant<int, 2>());
}

}  // extern "C"
