// wadges_device.hpp -- device-side building blocks shared by the WADGES kernels:
// per-node primitive records, the two-point entropy-conservative (Chandrashekar)
// flux in contracted form, and the entropy <-> conservative variable maps.
//
// Math follows the reference (ntan15/scratch_wadge):
//   logmean      okl/Euler3D.okl:55-83, okl/Euler2D.okl:54-66
//   flux         okl/Euler3D.okl:122-162, okl/Euler2D.okl:139-161
//   VU / UV      okl/Euler3D.okl:24-51,85-120, okl/Euler2D.okl:19-33,69-98
// restated for CDNA4 FP64:
//   * per-node quantities (velocity, beta = rho/2p, ln rho, ln beta, |u|^2/2)
//     are computed once per node instead of once per pair;
//   * the log branch of the log mean, (aL-aR)/ln(aL/aR), uses the per-node
//     logs: (aL-aR)/(ln aL - ln aR); the series branch keeps the reference's
//     thresholds (1e-4 with 4 terms in 3D, 1e-3 with 5 terms in 2D);
//   * the three divisions of a flux evaluation (two log means, the pressure
//     average and rho_log/beta_log) share one reciprocal;
//   * the flux is never materialised as a DIM x NFLD matrix: callers contract it
//     with a direction vector n (sum_d n_d F_d) through flux_dot().
// All arithmetic is FP64.
#pragma once

#include <hip/hip_runtime.h>

namespace wadges {

template <int DIM> struct Dims {
  static constexpr int NFLD = DIM + 2;   // conserved fields
  static constexpr int NG = DIM * DIM;   // geometric factors per node
  static constexpr int NFACES = DIM + 1; // faces per simplex
  static constexpr int NFG = DIM + 1;    // face geometric factors per face point
};

// Primitive record of a node: what the flux needs from one side of a pair.
template <int DIM> struct Node {
  double rho;
  double hv[DIM];  // half velocity, u/2
  double beta;     // rho / (2 p)
  double lrho;     // ln rho
  double lbeta;    // ln beta
  double hq;       // |u|^2 / 2
};

// 1/x from v_rcp_f64 refined by two Newton steps (the first half of the compiler's
// IEEE division sequence, without the final quotient correction): ~1 ulp.
__device__ __forceinline__ double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// Conservative state -> node record.  3D multiplies by 1/rho
// (okl/Euler3D.okl:2252-2256, TESTING); 2D divides (okl/Euler2D.okl:1127-1129).
template <int DIM>
__device__ __forceinline__ Node<DIM> make_node(const double (&U)[DIM + 2], double gm1) {
  Node<DIM> n;
  n.rho = U[0];
  double v[DIM];
  if constexpr (DIM == 3) {
    const double inv = 1.0 / U[0];
    for (int d = 0; d < DIM; ++d) v[d] = U[1 + d] * inv;
  } else {
    for (int d = 0; d < DIM; ++d) v[d] = U[1 + d] / U[0];
  }
  double uu = v[0] * v[0];
  for (int d = 1; d < DIM; ++d) uu = uu + v[d] * v[d];
  const double p = gm1 * (U[DIM + 1] - 0.5 * U[0] * uu);
  if constexpr (DIM == 3)
    n.beta = U[0] / (2.0 * p);
  else
    n.beta = 0.5 * U[0] / p;
  for (int d = 0; d < DIM; ++d) n.hv[d] = 0.5 * v[d];
  n.lrho = log(n.rho);
  n.lbeta = log(n.beta);
  n.hq = 0.5 * uu;
  return n;
}

// Pair quantities of the EC flux; F_d = flux_dot with n = e_d.
template <int DIM> struct PairFlux {
  double rholog, pa, f4aux;
  double va[DIM];  // averaged velocity
};

template <int DIM> __device__ __forceinline__ double logmean_series(double u) {
  // 1 + u/3 + u^2/5 + u^3/7 (+ u^4/9 in 2D), Horner form
  if constexpr (DIM == 3)
    return 1.0 + u * (1.0 / 3.0 + u * (1.0 / 5.0 + u * (1.0 / 7.0)));
  else
    return 1.0 + u * (1.0 / 3.0 + u * (1.0 / 5.0 + u * (1.0 / 7.0 + u * (1.0 / 9.0))));
}

// inv2gm1 = 1 / (2 (gamma - 1)).  Both log-mean branches are evaluated and
// selected without branching (lanes of a wave take different branches).
template <int DIM>
__device__ __forceinline__ PairFlux<DIM> pair_flux(const Node<DIM> &a, const Node<DIM> &b,
                                                   double inv2gm1) {
  constexpr double TH = (DIM == 3) ? 1.0e-4 : 1.0e-3;
  const double sr = a.rho + b.rho, dr = a.rho - b.rho;
  const double sb = a.beta + b.beta, db = a.beta - b.beta;
  // f = (aL-aR)/(aL+aR) for both log means from one reciprocal
  const double q1 = rcp(sr * sb);
  const double fr = dr * sb * q1;
  const double fb = db * sr * q1;
  const double ur = fr * fr, ub = fb * fb;
  const double polyR = 2.0 * logmean_series<DIM>(ur);
  const double polyB = 2.0 * logmean_series<DIM>(ub);
  const double ldR = a.lrho - b.lrho;
  const double ldB = a.lbeta - b.lbeta;
  const bool serR = ur < TH, serB = ub < TH;
  // logmean = num / den:  series (aL+aR) / (2 F(u)),  log (aL-aR) / (ln aL - ln aR)
  const double numR = serR ? sr : dr;
  const double denR = serR ? polyR : ldR;
  const double numB = serB ? sb : db;
  const double denB = serB ? polyB : ldB;
  // rholog = numR/denR, rholog/betalog = numR denB/(denR numB), pa = (sr/2)/sb
  const double q = rcp(denR * numB * sb);
  const double A = numB * q;
  PairFlux<DIM> P;
  P.rholog = numR * sb * A;
  P.pa = 0.5 * sr * denR * A;
  const double ratio = numR * denB * sb * q;
  double vv = 0.0;
  for (int d = 0; d < DIM; ++d) {
    P.va[d] = a.hv[d] + b.hv[d];
    vv = vv + P.va[d] * P.va[d];
  }
  const double vnavg = 2.0 * vv - (a.hq + b.hq);
  P.f4aux = ratio * inv2gm1 + P.pa + 0.5 * P.rholog * vnavg;
  return P;
}

// out[fld] = sum_d n_d F_d[fld] for the EC flux of a pair.
template <int DIM>
__device__ __forceinline__ void flux_dot(const PairFlux<DIM> &P, const double (&n)[DIM],
                                         double (&out)[DIM + 2]) {
  double Un = P.va[0] * n[0];
  for (int d = 1; d < DIM; ++d) Un = Un + P.va[d] * n[d];
  const double rUn = P.rholog * Un;
  out[0] = rUn;
  for (int d = 0; d < DIM; ++d) out[1 + d] = rUn * P.va[d] + P.pa * n[d];
  out[DIM + 1] = P.f4aux * Un;
}

// Conservative -> entropy variables.  s = ln((gamma-1) rho e) - gamma ln rho
// (the reference evaluates ln((gamma-1) rho e / rho^gamma)).
template <int DIM>
__device__ __forceinline__ void entropy_vars(const double (&U)[DIM + 2], double gamma,
                                             double (&V)[DIM + 2]) {
  const double rho = U[0], E = U[DIM + 1];
  double mm = U[1] * U[1];
  for (int d = 1; d < DIM; ++d) mm = mm + U[1 + d] * U[1 + d];
  const double rhoe = E - 0.5 * mm / rho;
  const double s = log((gamma - 1.0) * rhoe) - gamma * log(rho);
  const double invrhoe = 1.0 / rhoe;
  V[0] = (-E) * invrhoe + ((gamma + 1.0) - s);
  for (int d = 0; d < DIM; ++d) V[1 + d] = U[1 + d] * invrhoe;
  V[DIM + 1] = (-rho) * invrhoe;
}

// Entropy -> conservative variables.
// rho e = ((gamma-1)/(-v_last)^gamma)^(1/(gamma-1)) exp(-s/(gamma-1))
//       = exp((ln(gamma-1) - gamma ln(-v_last) - s) / (gamma-1)).
template <int DIM>
__device__ __forceinline__ void conservative_vars(const double (&V)[DIM + 2], double gamma,
                                                  double lgm1, double (&U)[DIM + 2]) {
  const double vl = V[DIM + 1];
  double vv = V[1] * V[1];
  for (int d = 1; d < DIM; ++d) vv = vv + V[1 + d] * V[1 + d];
  const double h = vv / (2.0 * vl);
  const double s = gamma - V[0] + h;
  const double rhoe = exp((lgm1 - gamma * log(-vl) - s) / (gamma - 1.0));
  U[0] = rhoe * (-vl);
  for (int d = 0; d < DIM; ++d) U[1 + d] = rhoe * V[1 + d];
  U[DIM + 1] = rhoe * (1.0 - h);
}

// Deterministic accumulation into an LDS accumulator.  The kernels schedule
// pairs so that within one instruction no two lanes of a wave add into the
// same address, and every element's lanes live in one wave; the order of the
// additions is therefore fixed by the program.
__device__ __forceinline__ void lds_add(double *p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

}  // namespace wadges
