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
//   * the divisions of a flux evaluation (two log means, the pressure average
//     and rho_log/beta_log) share two reciprocals, one on smooth states;
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
// LOGS = false leaves ln rho and ln beta out, for a node that serves a single
// pair: pair_flux<DIM, true> then takes the logarithms only where a log-mean
// log branch needs them.
template <int DIM, bool LOGS = true>
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
  n.lrho = LOGS ? log(n.rho) : 0.0;
  n.lbeta = LOGS ? log(n.beta) : 0.0;
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

// 2 * logmean_series(u), bit for bit: scaling by 2 is exact, so Horner with
// doubled constants rounds every step to twice the value (one multiply fewer).
template <int DIM> __device__ __forceinline__ double logmean_series2(double u) {
  if constexpr (DIM == 3)
    return 2.0 + u * (2.0 / 3.0 + u * (2.0 / 5.0 + u * (2.0 / 7.0)));
  else
    return 2.0 + u * (2.0 / 3.0 + u * (2.0 / 5.0 + u * (2.0 / 7.0 + u * (2.0 / 9.0))));
}

// 1 / (2 F(u)) for the 3D series F(u) = 1 + u/3 + u^2/5 + u^3/7, as the
// series of 1/F to the same order: 1 - u/3 - 4u^2/45 - 44u^3/945 (the
// coefficients c_n = -sum_k a_k c_{n-k} of the reciprocal of a power series).
// For u below the series threshold 1e-4 the omitted term is below 1e-17
// relative, under the rounding of the division it replaces.
__device__ __forceinline__ double logmean_inv_series2_3d(double u) {
  return 0.5 + u * (-1.0 / 6.0 + u * (-2.0 / 45.0 + u * (-22.0 / 945.0)));
}

// True when p holds on every active lane of the wave (a scalar test).
__device__ __forceinline__ bool wave_all(bool p) {
  return __builtin_amdgcn_ballot_w64(p) == __builtin_amdgcn_read_exec();
}

// A flag the caller knows to be the same on every lane, as a value the
// compiler knows to be uniform (a scalar register, tested with a scalar branch
// instead of lane masks).
__device__ __forceinline__ bool uniform_flag(bool p) {
  return __builtin_amdgcn_readfirstlane((int)p) != 0;
}

// Smooth waves (3D).  A node is in band when its rho and beta lie within a
// relative kSmoothBand of a reference node's.  When every node of a wave's
// elements is in band, every pair (a, b) of them has |a - b| / (a + b) <=
// band / (1 - band) < 0.0099, so f^2 < 1e-4: both log means of every pair take
// the 3D series branch, and the pair routines may skip their per-pair wave test
// (the branch they take is the same, so are the bits).
constexpr double kSmoothBand = 0.0098;

__device__ __forceinline__ bool in_smooth_band(double rho, double beta, double rho_ref,
                                               double beta_ref) {
  return fabs(rho - rho_ref) <= kSmoothBand * rho_ref &&
         fabs(beta - beta_ref) <= kSmoothBand * beta_ref;
}

// inv2gm1 = 1 / (2 (gamma - 1)).  Lanes of a wave may take different log-mean
// branches: both are evaluated and selected per lane, with two reciprocals per
// pair (f of both log means, then the quotients).  In 3D, when every active
// lane takes the series branch for both log means (smooth states: neighbouring
// nodes differ by far less than the 1 % that the series threshold allows), a
// wave-uniform test skips the selection.  With ONE_RCP that path also needs
// one reciprocal only: 1/(rho_a + rho_b) and 1/(beta_a + beta_b) both come
// from 1/((rho_a + rho_b)(beta_a + beta_b)), and rho_log = (rho_a + rho_b) /
// (2 F(u_rho)) uses the series of 1/(2F) instead of a division (equal to
// round-off); without it the skip gives the same bits as the selection.  The
// kernels whose register budget has room for the second path's values use
// ONE_RCP.  (In 2D the kernels aim at four or more waves per SIMD, whose
// register budget cannot hold both branches' values across the test: it
// spilled.)
// smooth: a wave-uniform flag by which the caller guarantees (3D) that every
// active lane's pair takes the series branch of both log means; the per-pair
// wave test is then skipped (the same branch, so the same bits).
template <int DIM, bool LAZY_LOGS = false, bool ONE_RCP = false>
__device__ __forceinline__ PairFlux<DIM> pair_flux(const Node<DIM> &a, const Node<DIM> &b,
                                                   double inv2gm1, bool smooth = false) {
  constexpr double TH = (DIM == 3) ? 1.0e-4 : 1.0e-3;
  const double sr = a.rho + b.rho, dr = a.rho - b.rho;
  const double sb = a.beta + b.beta, db = a.beta - b.beta;
  // f = (aL-aR)/(aL+aR) for both log means from one reciprocal
  const double q1 = rcp(sr * sb);
  const double fr = dr * sb * q1;
  const double fb = db * sr * q1;
  const double ur = fr * fr, ub = fb * fb;
  const bool serR = ur < TH, serB = ub < TH;
  PairFlux<DIM> P;
  double ratio;   // rho_log / beta_log
  const bool all_series = DIM == 3 && (smooth || wave_all(serR && serB));
  if (ONE_RCP && all_series) {
    const double isb = sr * q1;                       // 1 / sb
    P.rholog = sr * logmean_inv_series2_3d(ur);
    P.pa = 0.5 * sr * isb;
    ratio = P.rholog * logmean_series2<DIM>(ub) * isb;   // beta_log = sb / (2 F(u_beta))
  } else {
    const double polyR = logmean_series2<DIM>(ur);
    const double polyB = logmean_series2<DIM>(ub);
    // logmean = num / den:  series (aL+aR) / (2 F(u)),  log (aL-aR) / (ln aL - ln aR)
    double numR = sr, denR = polyR, numB = sb, denB = polyB;
    if (!all_series) {
      const double ldR = LAZY_LOGS ? log(a.rho) - log(b.rho) : a.lrho - b.lrho;
      const double ldB = LAZY_LOGS ? log(a.beta) - log(b.beta) : a.lbeta - b.lbeta;
      numR = serR ? sr : dr;
      denR = serR ? polyR : ldR;
      numB = serB ? sb : db;
      denB = serB ? polyB : ldB;
    }
    // rholog = numR/denR, rholog/betalog = numR denB/(denR numB), pa = (sr/2)/sb
    const double q = rcp(denR * numB * sb);
    const double A = numB * q;
    P.rholog = numR * sb * A;
    P.pa = 0.5 * sr * denR * A;
    ratio = numR * denB * sb * q;
  }
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

// lds_add for a second group of lanes whose addresses coincide with the first
// group's (taken under the complementary lane mask right after it).  Its
// atomic carries another scope, the wave's (the addresses lie in the calling
// wave's own LDS region, which no other wave touches), so the compiler cannot
// merge the two masked adds into one instruction in which two lanes would add
// into one address in an order the program does not fix.
__device__ __forceinline__ void lds_add_second(double *p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

}  // namespace wadges

// ======================================================================
// Reverse mode (discrete adjoint) building blocks.  Hand-written VJPs of the
// forward routines above; derivatives are taken of the mathematical functions
// the forward kernels evaluate (the log mean through whichever branch the
// forward takes, max(lamM, lamP) through its larger argument, M at ties).
// ======================================================================

namespace wadges {

// Cotangent of a node's primitive quantities (rho, velocity u, beta).
template <int DIM> struct PrimBar {
  double rho, u[DIM], beta;
};

template <int DIM> __device__ __forceinline__ void primbar_zero(PrimBar<DIM> &g) {
  g.rho = 0.0;
  for (int d = 0; d < DIM; ++d) g.u[d] = 0.0;
  g.beta = 0.0;
}

// Partial derivatives (dl/da, dl/db) of the log mean l(a, b), in the branch the
// forward evaluation uses: series l = (a+b)/(2F(f^2)), f = (a-b)/(a+b), or
// l = (a-b)/(ln a - ln b).
template <int DIM>
__device__ __forceinline__ void logmean_partials(double a, double b, double la, double lb,
                                                 double &l, double &da, double &db) {
  constexpr double TH = (DIM == 3) ? 1.0e-4 : 1.0e-3;
  const double s = a + b, d = a - b;
  const double is = rcp(s);
  const double f = d * is;
  const double u = f * f;
  if (u < TH) {
    const double F = logmean_series<DIM>(u);
    double Fp;  // F'(u)
    if constexpr (DIM == 3)
      Fp = 1.0 / 3.0 + u * (2.0 / 5.0 + u * (3.0 / 7.0));
    else
      Fp = 1.0 / 3.0 + u * (2.0 / 5.0 + u * (3.0 / 7.0 + u * (4.0 / 9.0)));
    const double iF = rcp(F);
    l = 0.5 * s * iF;
    const double t = 2.0 * f * Fp * is * iF * iF;
    da = 0.5 * iF - t * b;
    db = 0.5 * iF + t * a;
  } else {
    const double iL = rcp(la - lb);
    l = d * iL;
    da = (1.0 - l * rcp(a)) * iL;
    db = (l * rcp(b) - 1.0) * iL;
  }
}

// VJP of L = sum_d sum_fld Gb[d][fld] F_d[fld](a, b) with respect to the
// primitive quantities of both nodes, accumulated into ga, gb, for a cotangent
// of rank NT, Gb[d][fld] = sum_k n_k[d] c_k[fld] (rank one: the contracted
// volume-face and surface couplings; rank two: a volume pair's own row and its
// partner row).  L depends on the flux through A = sum_d Gb[d][0] va_d,
// B = sum_dm Gb[d][1+m] va_d va_m, C = sum_d Gb[d][last] va_d and
// tr = sum_d Gb[d][1+d]; with Un_k = n_k.va and cm_k = sum_m c_k[1+m] va_m these
// are A = sum c_k[0] Un_k, B = sum Un_k cm_k, C = sum c_k[last] Un_k,
// tr = sum_k sum_d n_k[d] c_k[1+d], and the velocity derivative needs, per
// direction, t_d = sum_k n_k[d] (c_k[0] + cm_k) + c_k[1+d] Un_k:
// O(NT (DIM + NFLD)) work instead of the O(DIM^2 NFLD) of a dense Gb.
template <int DIM, int NT>
__device__ __forceinline__ void flux_dot_vjp_k(const Node<DIM> &a, const Node<DIM> &b,
                                               const double (&n)[NT][DIM],
                                               const double (&c)[NT][DIM + 2], double inv2gm1,
                                               PrimBar<DIM> &ga, PrimBar<DIM> &gb) {
  constexpr int NF = DIM + 2;
  double rl, rla, rlb, bl, bla, blb, isb, ibl;
  const double sr = a.rho + b.rho, sb = a.beta + b.beta;
  constexpr double TH = (DIM == 3) ? 1.0e-4 : 1.0e-3;
  const double dr = a.rho - b.rho, db = a.beta - b.beta;
  bool series = false;
  if constexpr (DIM == 3) {
    // f of both log means from one reciprocal; when every active lane takes
    // the series branch of both (smooth states), 1/F comes from the series of
    // 1/F and every other reciprocal of the pair from that one.
    const double q = rcp(sr * sb);
    const double isr = sb * q;
    const double isb1 = sr * q;
    const double fr = dr * isr, fb = db * isb1;
    const double ur = fr * fr, ub = fb * fb;
    series = wave_all(ur < TH && ub < TH);
    if (series) {
      // l = s/(2F(u)); dl/da = 1/(2F) - l' b, dl/db = 1/(2F) + l' a with
      // l' = 2 f F'(u) / (s F^2)  (the partials of the series branch)
      auto partials = [&](double sa, double is, double f, double u, double x, double y,
                          double &l, double &dx, double &dy, double &i2F) {
        const double h = logmean_inv_series2_3d(u);          // 1/(2F)
        const double Fp = 1.0 / 3.0 + u * (2.0 / 5.0 + u * (3.0 / 7.0));
        l = sa * h;
        const double t = 8.0 * f * Fp * is * h * h;           // 2 f F' / (s F^2)
        dx = h - t * y;
        dy = h + t * x;
        i2F = h;
      };
      double hR, hB;
      partials(sr, isr, fr, ur, a.rho, b.rho, rl, rla, rlb, hR);
      partials(sb, isb1, fb, ub, a.beta, b.beta, bl, bla, blb, hB);
      isb = isb1;
      ibl = isb1 * logmean_series2<DIM>(ub);                  // 1/bl = 2F(u_beta)/sb
      (void)hR;
      (void)hB;
    }
  }
  if (!series) {
    logmean_partials<DIM>(a.rho, b.rho, a.lrho, b.lrho, rl, rla, rlb);
    logmean_partials<DIM>(a.beta, b.beta, a.lbeta, b.lbeta, bl, bla, blb);
    isb = rcp(sb);
  }
  const double pa = 0.5 * sr * isb;
  double va[DIM], vv = 0.0;
  for (int d = 0; d < DIM; ++d) {
    va[d] = a.hv[d] + b.hv[d];
    vv = vv + va[d] * va[d];
  }
  const double vnavg = 2.0 * vv - (a.hq + b.hq);
  if (!series) ibl = rcp(bl);
  const double f4 = rl * inv2gm1 * ibl + pa + 0.5 * rl * vnavg;
  double Un[NT], c0m[NT], A = 0.0, B = 0.0, C = 0.0, tr = 0.0;
  for (int k = 0; k < NT; ++k) {
    double u = n[k][0] * va[0], cm = c[k][1] * va[0];
    for (int d = 1; d < DIM; ++d) {
      u = u + n[k][d] * va[d];
      cm = cm + c[k][1 + d] * va[d];
    }
    for (int d = 0; d < DIM; ++d) tr = tr + n[k][d] * c[k][1 + d];
    Un[k] = u;
    c0m[k] = c[k][0] + cm;
    A = A + c[k][0] * u;
    B = B + u * cm;
    C = C + c[k][NF - 1] * u;
  }
  const double L_rl = A + B + C * (inv2gm1 * ibl + 0.5 * vnavg);
  const double L_bl = -C * rl * inv2gm1 * ibl * ibl;
  const double L_pa = tr + C;
  for (int d = 0; d < DIM; ++d) {
    double t = 0.0, e = 0.0;
    for (int k = 0; k < NT; ++k) {
      t = t + n[k][d] * c0m[k] + c[k][1 + d] * Un[k];
      e = e + n[k][d] * c[k][NF - 1];
    }
    const double L_va = rl * t + f4 * e + C * 2.0 * rl * va[d];
    const double ua = 2.0 * a.hv[d], ub = 2.0 * b.hv[d];
    ga.u[d] += 0.5 * L_va - 0.5 * C * rl * ua;
    gb.u[d] += 0.5 * L_va - 0.5 * C * rl * ub;
  }
  ga.rho += L_rl * rla + L_pa * 0.5 * isb;
  gb.rho += L_rl * rlb + L_pa * 0.5 * isb;
  ga.beta += L_bl * bla - L_pa * pa * isb;
  gb.beta += L_bl * blb - L_pa * pa * isb;
}

template <int DIM>
__device__ __forceinline__ void flux_dot_vjp(const Node<DIM> &a, const Node<DIM> &b,
                                             const double (&n)[DIM], const double (&c)[DIM + 2],
                                             double inv2gm1, PrimBar<DIM> &ga, PrimBar<DIM> &gb) {
  double nk[1][DIM], ck[1][DIM + 2];
  for (int d = 0; d < DIM; ++d) nk[0][d] = n[d];
  for (int f = 0; f < DIM + 2; ++f) ck[0][f] = c[f];
  flux_dot_vjp_k<DIM, 1>(a, b, nk, ck, inv2gm1, ga, gb);
}

// Primitive cotangent -> conservative cotangent of one node:
// u = m / rho, beta = rho / (2p), p = (gamma - 1)(E - |m|^2/(2 rho)).
template <int DIM>
__device__ __forceinline__ void primbar_to_cons(const Node<DIM> &n, const PrimBar<DIM> &g,
                                                double gm1, double (&Ub)[DIM + 2]) {
  const double irho = rcp(n.rho);
  const double p = 0.5 * n.rho * rcp(n.beta);      // p = rho / (2 beta)
  const double bp = n.beta * rcp(p);                // beta / p
  double gu_u = 0.0;
  for (int d = 0; d < DIM; ++d) gu_u = gu_u + g.u[d] * (2.0 * n.hv[d]);
  // d beta/d rho = beta/rho - (beta/p)(gamma-1)|u|^2/2
  Ub[0] = g.rho - gu_u * irho + g.beta * (n.beta * irho - bp * gm1 * n.hq);
  for (int d = 0; d < DIM; ++d) Ub[1 + d] = g.u[d] * irho + g.beta * bp * gm1 * (2.0 * n.hv[d]);
  Ub[DIM + 1] = -g.beta * bp * gm1;
}

// Wave speed lambda = |u| + sqrt(gamma p / rho) and its gradient with respect
// to the conservative state.
template <int DIM>
__device__ __forceinline__ double wavespeed(const double (&U)[DIM + 2], double gamma, double gm1,
                                            double (&dl)[DIM + 2]) {
  const double irho = rcp(U[0]);
  double v[DIM], uu = 0.0;
  for (int d = 0; d < DIM; ++d) {
    v[d] = U[1 + d] * irho;
    uu = uu + v[d] * v[d];
  }
  const double p = gm1 * (U[DIM + 1] - 0.5 * U[0] * uu);
  const double su = sqrt(uu);
  const double c = sqrt(gamma * p * irho);
  // d|u|: d(|u|)/du_d = u_d/|u| (0 at rest); u_d = m_d/rho
  const double isu = su > 0.0 ? rcp(su) : 0.0;
  // c^2 = gamma p / rho: d(c)/dq = (gamma/(2c)) d(p/rho)/dq
  const double k = 0.5 * gamma * rcp(c);
  // p/rho = gm1 (E/rho - |m|^2/(2 rho^2))
  const double dpr_rho = gm1 * (-U[DIM + 1] * irho * irho + uu * irho);
  dl[0] = -uu * isu * irho + k * dpr_rho;
  for (int d = 0; d < DIM; ++d) dl[1 + d] = v[d] * isu * irho + k * (-gm1 * v[d] * irho);
  dl[DIM + 1] = k * gm1 * irho;
  return su + c;
}

// VJP of entropy_vars (U -> V) at U.
template <int DIM>
__device__ __forceinline__ void entropy_vars_vjp(const double (&U)[DIM + 2], double gamma,
                                                 const double (&Vb)[DIM + 2], double (&Ub)[DIM + 2]) {
  constexpr int NF = DIM + 2;
  const double rho = U[0], E = U[NF - 1];
  double mm = 0.0;
  for (int d = 0; d < DIM; ++d) mm = mm + U[1 + d] * U[1 + d];
  const double irho = rcp(rho);
  const double rhoe = E - 0.5 * mm * irho;
  const double io = rcp(rhoe);
  // V0 = -E io + (gamma + 1) - s ; V_m = m io ; V_last = -rho io
  double iob = Vb[0] * (-E) + Vb[NF - 1] * (-rho);
  for (int d = 0; d < DIM; ++d) iob = iob + Vb[1 + d] * U[1 + d];
  const double sb = -Vb[0];
  double rhob = -Vb[NF - 1] * io - sb * gamma * irho;     // s = ln((g-1) rho e) - g ln rho
  double Eb = -Vb[0] * io;
  const double rhoeb = -iob * io * io + sb * io;
  Eb = Eb + rhoeb;
  rhob = rhob + rhoeb * 0.5 * mm * irho * irho;
  Ub[0] = rhob;
  for (int d = 0; d < DIM; ++d) Ub[1 + d] = Vb[1 + d] * io - rhoeb * U[1 + d] * irho;
  Ub[NF - 1] = Eb;
}

// VJP of conservative_vars (V -> U) at V.
template <int DIM>
__device__ __forceinline__ void conservative_vars_vjp(const double (&V)[DIM + 2], double gamma,
                                                      double lgm1, const double (&Ub)[DIM + 2],
                                                      double (&Vb)[DIM + 2]) {
  constexpr int NF = DIM + 2;
  const double vl = V[NF - 1];
  double vv = 0.0;
  for (int d = 0; d < DIM; ++d) vv = vv + V[1 + d] * V[1 + d];
  const double ivl = rcp(vl);
  const double h = 0.5 * vv * ivl;
  const double s = gamma - V[0] + h;
  const double igm1 = 1.0 / (gamma - 1.0);
  const double rhoe = exp((lgm1 - gamma * log(-vl) - s) * igm1);
  double rhoeb = Ub[0] * (-vl) + Ub[NF - 1] * (1.0 - h);
  for (int d = 0; d < DIM; ++d) rhoeb = rhoeb + Ub[1 + d] * V[1 + d];
  double vlb = -Ub[0] * rhoe;
  double hb = -Ub[NF - 1] * rhoe;
  const double Xb = rhoeb * rhoe;
  const double sb = -Xb * igm1;
  vlb = vlb - Xb * gamma * igm1 * ivl;
  Vb[0] = -sb;
  hb = hb + sb;
  const double vvb = 0.5 * hb * ivl;
  vlb = vlb - hb * h * ivl;
  for (int d = 0; d < DIM; ++d) Vb[1 + d] = Ub[1 + d] * rhoe + 2.0 * vvb * V[1 + d];
  Vb[NF - 1] = vlb;
}

}  // namespace wadges
