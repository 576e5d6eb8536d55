// wadges_adjoint.hpp -- discrete adjoint (vector-Jacobian products) of the
// WADGES stage, included by wadges.hip.
//
// Forward stage (okl/Euler3D.okl:2123-2884, euler.c:3656-3666):
//   R   = vol(u~q, u~f) + lift(surf(u~f, u~f[nbr]) + rhsQf)      (RHS)
//   r   = VqPq diag(1/J) VqPq (-2 R)
//   res'= fa res + dt r ;  Q' = Q + fb res'
//   v~  = VqPq diag(1/J) VqPq (J v(Q')) ; u~q' = u(VqPq v~), u~f' = u(VfPq v~)
// The VJPs below run that sequence backwards:
//   sbar_kernel      s_bar = VqLq^T R_bar per element (the lift transposed); the
//                    RHS VJP needs it at neighbour face points, hence a pass of
//                    its own
//   rhs_vjp_kernel   (u~q_bar, u~f_bar) from R_bar: the flux-differencing pairs on
//                    the same diagonal schedules as the forward volume kernel (each
//                    symmetric pair differentiated once, partner cotangents
//                    accumulated in LDS), then the surface flux and LF penalty of
//                    this element's face points and of the neighbour's matching
//                    points (recomputed locally, so no scatter between elements)
//   update_vjp_kernel (Q_bar, res_bar, R_bar) from the stage outputs' cotangents,
//                    recomputing the projection from Q'
//   proj_vjp_kernel  the initial-state projection u~q = u(VqPq v(Q0)),
//                    u~f = u(VfPq v(Q0)) (euler.c:4028-4133)
// Included inside namespace wadges, after the forward kernels it reuses.
#pragma once

template <class S, int TS, int TPW>
__global__ void __launch_bounds__(WAVE * 4)
sbar_kernel(int K, const double *__restrict__ VqLq, const double *__restrict__ Rb,
            double *__restrict__ sbar) {
  constexpr int NQ = S::NQ, NFN = S::NFN, NFLD = S::NFLD;
  constexpr int EPB = 4 * TPW;
  __shared__ double sR[EPB][NFLD][NQ];
  const int wave = threadIdx.x / WAVE;
  const int wl = threadIdx.x % WAVE;
  const int tw = wl / TS;
  const int lane = wl - tw * TS;
  const bool active = tw < TPW;
  const int e = wave * TPW + (active ? tw : 0);
  const int k = blockIdx.x * EPB + e;
  const bool valid = active && k < K;
  if (valid && lane < NQ)
    for (int f = 0; f < NFLD; ++f) sR[e][f][lane] = Rb[(size_t)k * NQ * NFLD + f * NQ + lane];
  __syncthreads();
  if (valid && lane < NFN) {
    double v[NFLD];
    for (int f = 0; f < NFLD; ++f) v[f] = 0.0;
#pragma unroll 4
    for (int i = 0; i < NQ; ++i) {
      const double L = ld_off(VqLq, (unsigned)(i + lane * NQ));
      for (int f = 0; f < NFLD; ++f) v[f] += L * sR[e][f][i];
    }
    for (int f = 0; f < NFLD; ++f) sbar[(size_t)k * NFN * NFLD + f * NFN + lane] = v[f];
  }
}

template <class S> struct VjpSmem {
  static constexpr int NPB = S::DIM + 2;  // primitive cotangent components
  static constexpr int acc_v = NPB * S::NQ;
  static constexpr int phase1 = S::VREC * S::NQ + S::NFLD * S::NQ;
  static constexpr int phase2 = (S::FREC + S::NFLD + NPB) * S::NFN;
  static constexpr int per_elem = acc_v + (phase1 > phase2 ? phase1 : phase2);
};

template <int DIM>
__device__ __forceinline__ void lds_add_primbar(double *acc, int stride, int j,
                                                const PrimBar<DIM> &g) {
  lds_add(&acc[0 * stride + j], g.rho);
  for (int d = 0; d < DIM; ++d) lds_add(&acc[(1 + d) * stride + j], g.u[d]);
  lds_add(&acc[(DIM + 1) * stride + j], g.beta);
}

template <int DIM>
__device__ __forceinline__ void add_lds_primbar(PrimBar<DIM> &g, const double *acc, int stride,
                                                int j) {
  g.rho += acc[0 * stride + j];
  for (int d = 0; d < DIM; ++d) g.u[d] += acc[(1 + d) * stride + j];
  g.beta += acc[(DIM + 1) * stride + j];
}

template <class S, int TPW>
__global__ void __launch_bounds__(WAVE)
rhs_vjp_kernel(int K, const double *__restrict__ vgeo, const double *__restrict__ vfgeo,
               const double *__restrict__ fgeo, const uint32_t *__restrict__ mapPq,
               const double *__restrict__ nrJ, const double *__restrict__ nsJ,
               const double *__restrict__ ntJ, const double *__restrict__ ws,
               const double *__restrict__ Uq, const double *__restrict__ Uf,
               const double *__restrict__ Rb, const double *__restrict__ sbar,
               double *__restrict__ uqb, double *__restrict__ ufb, double gamma, double gm1,
               double inv2gm1, double ltau) {
  constexpr int DIM = S::DIM, NQ = S::NQ, NFQ = S::NFQ, NFN = S::NFN, NFLD = S::NFLD;
  constexpr int NG = S::NG, TVT = S::TVT, VS = S::VS, VREC = S::VREC, FREC = S::FREC;
  constexpr int NPB = VjpSmem<S>::NPB, PER = VjpSmem<S>::per_elem, NFG = DIM + 1;
  constexpr int T = vf_steps<S>();
  extern __shared__ double smem[];
  __shared__ double s_nref[S::NFACES][DIM];

  const int lane = threadIdx.x;
  const int team = lane / NQ;
  const int i = lane - team * NQ;
  const bool active = team < TPW;
  const int k = blockIdx.x * TPW + team;
  const bool valid = active && k < K;
  const size_t kk = valid ? (size_t)k : 0;

  load_nref<S>(lane, nrJ, nsJ, ntJ, s_nref);
  double *gV = smem + (active ? team : 0) * PER;   // [NPB][NQ] partner cotangents
  double *vrec = gV + NPB * NQ;                    // phase 1
  double *RbV = vrec + VREC * NQ;                  // phase 1: R_bar rows [NFLD][NQ]
  double *frec = vrec;                             // phase 2
  double *sbF = frec + FREC * NFN;                 // phase 2: s_bar own [NFLD][NFN]
  double *gF = sbF + NFLD * NFN;                   // phase 2: [NPB][NFN]

  // ---------------- phase 1: records, volume-volume pairs
  Node<DIM> me;
  double hG[NG], Rbi[NFLD];
  PrimBar<DIM> gI;
  primbar_zero<DIM>(gI);
  if (valid) {
    double U[NFLD];
    for (int f = 0; f < NFLD; ++f) U[f] = Uq[kk * NQ * NFLD + f * NQ + i];
    me = make_node<DIM>(U, gm1);
    for (int c = 0; c < NG; ++c) hG[c] = 0.5 * vgeo[kk * NQ * NG + c * NQ + i];
    store_node<DIM>(vrec, NQ, i, me);
    for (int c = 0; c < NG; ++c) vrec[(VS + c) * NQ + i] = hG[c];
    for (int f = 0; f < NFLD; ++f) {
      Rbi[f] = Rb[kk * NQ * NFLD + f * NQ + i];
      RbV[f * NQ + i] = Rbi[f];
    }
    for (int c = 0; c < NPB; ++c) gV[c * NQ + i] = 0.0;
  }
  __syncthreads();
  if (valid) {
    const double *Dfw = ws;
    const double *Dbw = ws + DIM * TVT * NQ;
    for (int t = 0; t < TVT; ++t) {
      if (2 * t > NQ) break;
      if (2 * t == NQ && i >= NQ / 2) continue;
      int j = i + t;
      if (j >= NQ) j -= NQ;
      const Node<DIM> pj = load_node<DIM>(vrec, NQ, j);
      double Ga[NG];
      for (int c = 0; c < NG; ++c) Ga[c] = hG[c] + vrec[(VS + c) * NQ + j];
      double Dxf[DIM], Dxb[DIM];
      for (int d = 0; d < DIM; ++d) Dxf[d] = Dxb[d] = 0.0;
      for (int r = 0; r < DIM; ++r) {
        const double a = ld_off(Dfw, (unsigned)((r * TVT + t) * NQ + i));
        const double b = ld_off(Dbw, (unsigned)((r * TVT + t) * NQ + i));
        for (int d = 0; d < DIM; ++d) {
          Dxf[d] = Dxf[d] + a * Ga[r * DIM + d];
          Dxb[d] = Dxb[d] + b * Ga[r * DIM + d];
        }
      }
      // cotangent Rb_i (x) Dxf + Rb_j (x) Dxb (the partner row only off the diagonal)
      PrimBar<DIM> gj;
      primbar_zero<DIM>(gj);
      if (t > 0) {
        double nk[2][DIM], ck[2][NFLD];
        for (int d = 0; d < DIM; ++d) {
          nk[0][d] = Dxf[d];
          nk[1][d] = Dxb[d];
        }
        for (int f = 0; f < NFLD; ++f) {
          ck[0][f] = Rbi[f];
          ck[1][f] = RbV[f * NQ + j];
        }
        flux_dot_vjp_k<DIM, 2>(me, pj, nk, ck, inv2gm1, gI, gj);
      } else {
        flux_dot_vjp<DIM>(me, pj, Dxf, Rbi, inv2gm1, gI, gj);
      }
      if (t == 0) {
        gI.rho += gj.rho;
        for (int d = 0; d < DIM; ++d) gI.u[d] += gj.u[d];
        gI.beta += gj.beta;
      } else {
        lds_add_primbar<DIM>(gV, NQ, j, gj);
      }
    }
  }
  __syncthreads();

  // ---------------- phase 2: face records, s_bar, volume-face pairs.  Face
  // point work of all teams is spread over the whole wave.
  auto team_base = [&](int t) { return smem + t * PER; };
  constexpr int OFF_FREC = NPB * NQ, OFF_SBF = OFF_FREC + FREC * NFN, OFF_GF = OFF_SBF + NFLD * NFN;
  for (int p = lane; p < TPW * NFN; p += WAVE) {
    const int t = p / NFN;
    const int fn = p - t * NFN;
    const size_t kt = (size_t)blockIdx.x * TPW + t;
    if (kt >= (size_t)K) continue;
    double *frt = team_base(t) + OFF_FREC;
    double *sbt = team_base(t) + OFF_SBF;
    double *gft = team_base(t) + OFF_GF;
    double U[NFLD];
    for (int f = 0; f < NFLD; ++f) U[f] = Uf[kt * NFN * NFLD + f * NFN + fn];
    store_node<DIM>(frt, NFN, fn, make_node<DIM>(U, gm1));
    const int face = fn / NFQ;
    for (int d = 0; d < DIM; ++d) {
      double b = 0.0;
      for (int r = 0; r < DIM; ++r)
        b = b + 0.5 * vfgeo[kt * NFN * NG + (r * DIM + d) * NFN + fn] * s_nref[face][r];
      frt[(VS + d) * NFN + fn] = b;
    }
    for (int f = 0; f < NFLD; ++f) sbt[f * NFN + fn] = sbar[kt * NFN * NFLD + f * NFN + fn];
    for (int c = 0; c < NPB; ++c) gft[c * NFN + fn] = 0.0;
  }
  __syncthreads();
  if (valid) {
    const double *W1 = ws + S::WS_D;
    const double *W2 = W1 + T * NQ;
    auto vf_pair = [&](int t, int fn, const double (&a)[DIM]) {
      const Node<DIM> pf = load_node<DIM>(frec, NFN, fn);
      double n[DIM];
      for (int d = 0; d < DIM; ++d) n[d] = a[d] + frec[(VS + d) * NFN + fn];
      const double w1 = ld_off(W1, (unsigned)(t * NQ + i));
      const double w2 = ld_off(W2, (unsigned)(t * NQ + i));
      double c[NFLD];
      for (int f = 0; f < NFLD; ++f) c[f] = w1 * Rbi[f] + w2 * sbF[f * NFN + fn];
      PrimBar<DIM> gf;
      primbar_zero<DIM>(gf);
      flux_dot_vjp<DIM>(me, pf, n, c, inv2gm1, gI, gf);
      lds_add_primbar<DIM>(gF, NFN, fn, gf);
    };
    auto own_normal = [&](int face, double (&a)[DIM]) {
      for (int d = 0; d < DIM; ++d) {
        double v = 0.0;
        for (int r = 0; r < DIM; ++r) v = v + hG[r * DIM + d] * s_nref[face][r];
        a[d] = v;
      }
    };
    if constexpr (face_chunked<S>()) {
      const int g = i / NFQ, o = i - (i / NFQ) * NFQ;
      for (int c = 0; c < S::NFACES; ++c) {
        int face = c + g;
        if (face >= S::NFACES) face -= S::NFACES;
        double a[DIM];
        own_normal(face, a);
        for (int m = 0; m < NFQ; ++m) {
          int node = o + m;
          if (node >= NFQ) node -= NFQ;
          vf_pair(c * NFQ + m, face * NFQ + node, a);
        }
      }
    } else {
      for (int t = 0; t < T; ++t) {
        const int fn = vf_partner<S>(t, i);
        if (fn >= NFN) continue;
        double a[DIM];
        own_normal(fn / NFQ, a);
        vf_pair(t, fn, a);
      }
    }
  }
  __syncthreads();

  // ---------------- phase 3: surface terms at the face points, spread over the
  // whole wave.  A point's conservative-space part overwrites its s_bar slot
  // (only this lane reads that slot from here on).
  for (int p = lane; p < TPW * NFN; p += WAVE) {
    const int t = p / NFN;
    const int fn = p - t * NFN;
    const size_t kt = (size_t)blockIdx.x * TPW + t;
    if (kt >= (size_t)K) continue;
    double *sbt = team_base(t) + OFF_SBF;
    double *gft = team_base(t) + OFF_GF;
    const double *g = fgeo + kt * NFN * NFG;
    const size_t idM = (size_t)fn + kt * NFN * NFLD;
    const size_t idP = mapPq[kt * NFN + fn];
    double UM[NFLD], UP[NFLD], sbM[NFLD], cons[NFLD];
    for (int f = 0; f < NFLD; ++f) {
      UM[f] = Uf[idM + (size_t)f * NFN];
      UP[f] = Uf[idP + (size_t)f * NFN];
      sbM[f] = sbt[f * NFN + fn];
      cons[f] = 0.0;
    }
    const Node<DIM> a = make_node<DIM>(UM, gm1);
    const Node<DIM> b = make_node<DIM>(UP, gm1);
    double hnM[DIM];
    for (int d = 0; d < DIM; ++d) hnM[d] = 0.5 * g[fn + d * NFN];
    const double sJM = g[fn + DIM * NFN];
    PrimBar<DIM> ga, gb;
    primbar_zero<DIM>(ga);
    primbar_zero<DIM>(gb);
    flux_dot_vjp<DIM>(a, b, hnM, sbM, inv2gm1, ga, gb);
    double dlM[NFLD], dlP[NFLD];
    const double lamM = wavespeed<DIM>(UM, gamma, gm1, dlM);
    const double lamP = wavespeed<DIM>(UP, gamma, gm1, dlP);
    if (idP == idM) {
      // no neighbour: u~P = u~M, zero jump; both flux arguments are this point
      ga.rho += gb.rho;
      for (int d = 0; d < DIM; ++d) ga.u[d] += gb.u[d];
      ga.beta += gb.beta;
    } else {
      // own LF term  -ltau max(lamM, lamP) sJM (UP - UM) . sbM
      const double LfM = ltau * fmax(lamM, lamP) * sJM;
      double S1 = 0.0;
      for (int f = 0; f < NFLD; ++f) S1 = S1 + sbM[f] * (UP[f] - UM[f]);
      const double dM = (lamM >= lamP) ? -S1 * ltau * sJM : 0.0;
      for (int f = 0; f < NFLD; ++f) cons[f] += LfM * sbM[f] + dM * dlM[f];
      // the neighbour's terms at its matching point, where this point is the
      // second flux argument and the P side of its penalty
      const size_t eN = idP / ((size_t)NFN * NFLD);
      const int pN = (int)(idP - eN * (size_t)NFN * NFLD);
      const double *gN = fgeo + eN * NFN * NFG;
      double hnP[DIM], sbP[NFLD];
      for (int d = 0; d < DIM; ++d) hnP[d] = 0.5 * gN[pN + d * NFN];
      const double sJP = gN[pN + DIM * NFN];
      for (int f = 0; f < NFLD; ++f) sbP[f] = sbar[idP + (size_t)f * NFN];
      PrimBar<DIM> gx;
      primbar_zero<DIM>(gx);
      flux_dot_vjp<DIM>(b, a, hnP, sbP, inv2gm1, gx, ga);
      const double LfP = ltau * fmax(lamP, lamM) * sJP;
      double S2 = 0.0;
      for (int f = 0; f < NFLD; ++f) S2 = S2 + sbP[f] * (UM[f] - UP[f]);
      const double dP = (lamM > lamP) ? -S2 * ltau * sJP : 0.0;
      for (int f = 0; f < NFLD; ++f) cons[f] += -LfP * sbP[f] + dP * dlM[f];
    }
    // plain adds: this lane owns face point (t, fn) after the barrier
    gft[0 * NFN + fn] += ga.rho;
    for (int d = 0; d < DIM; ++d) gft[(1 + d) * NFN + fn] += ga.u[d];
    gft[(DIM + 1) * NFN + fn] += ga.beta;
    for (int f = 0; f < NFLD; ++f) sbt[f * NFN + fn] = cons[f];
  }
  __syncthreads();

  // ---------------- phase 4: primitive -> conservative cotangents
  if (valid) {
    add_lds_primbar<DIM>(gI, gV, NQ, i);
    double Ub[NFLD];
    primbar_to_cons<DIM>(me, gI, gm1, Ub);
    for (int f = 0; f < NFLD; ++f) uqb[kk * NQ * NFLD + f * NQ + i] = Ub[f];
  }
  for (int p = lane; p < TPW * NFN; p += WAVE) {
    const int t = p / NFN;
    const int fn = p - t * NFN;
    const size_t kt = (size_t)blockIdx.x * TPW + t;
    if (kt >= (size_t)K) continue;
    const double *frt = team_base(t) + OFF_FREC;
    const double *sbt = team_base(t) + OFF_SBF;
    double *gft = team_base(t) + OFF_GF;
    const Node<DIM> nf = load_node<DIM>(frt, NFN, fn);
    PrimBar<DIM> gf;
    primbar_zero<DIM>(gf);
    add_lds_primbar<DIM>(gf, gft, NFN, fn);
    double Ub[NFLD];
    primbar_to_cons<DIM>(nf, gf, gm1, Ub);
    for (int f = 0; f < NFLD; ++f) ufb[kt * NFN * NFLD + f * NFN + fn] = Ub[f] + sbt[f * NFN + fn];
  }
}

// (Q_bar, res_bar) <- stage VJP; R_bar written.  Q is the stage's output state.
// Workgroup shape of the update VJP.  2D: one wave (its barriers cost nothing),
// the small operators read through L1.  3D: four waves sharing the operators
// staged in LDS (the transposed products read them along rows, which through
// L1 touches a cache line per lane).  3D N = 3 declares three waves per SIMD,
// as many as its LDS allows (43 KB per four-wave workgroup).
template <class S> constexpr int uvjp_wg_waves() { return S::DIM == 2 ? 1 : 4; }
template <class S> constexpr int update_vjp_waves() { return (S::DIM == 3 && S::NQ == 23) ? 3 : 1; }

template <class S, int TPW>
__global__ void __launch_bounds__(WAVE * uvjp_wg_waves<S>()) __attribute__((amdgpu_waves_per_eu(update_vjp_waves<S>())))
update_vjp_kernel(int K, const double *__restrict__ Jq, const double *__restrict__ VqPq,
                  const double *__restrict__ VfPq, double fa, double fb, double fdt,
                  const double *__restrict__ Q, double *__restrict__ Qb,
                  double *__restrict__ resb, const double *__restrict__ uqb,
                  const double *__restrict__ ufb, double *__restrict__ Rb, double gamma,
                  double lgm1, const double *__restrict__ VqLq, double *__restrict__ sbar) {
  constexpr int DIM = S::DIM, NQ = S::NQ, NFN = S::NFN, NFLD = S::NFLD;
  constexpr int WPB = uvjp_wg_waves<S>();
  constexpr bool OPS_LDS = WPB > 1;
  constexpr int EPB = WPB * TPW;
  __shared__ double sP[OPS_LDS ? NQ * NQ : 1];
  __shared__ double sPf[OPS_LDS ? NFN * NQ : 1];
  __shared__ double bufA[EPB][NFLD][NQ];
  __shared__ double bufB[EPB][NFLD][NQ];
  __shared__ double bufF[EPB][NFLD][NFN];
  if constexpr (OPS_LDS) {
    for (int idx = threadIdx.x; idx < NQ * NQ; idx += blockDim.x) sP[idx] = VqPq[idx];
    for (int idx = threadIdx.x; idx < NFN * NQ; idx += blockDim.x) sPf[idx] = VfPq[idx];
  }
  auto Pv = [&](int idx) {      // VqPq, column-major
    if constexpr (OPS_LDS) return sP[idx]; else return ld_off(VqPq, (unsigned)idx);
  };
  auto Pf = [&](int idx) {      // VfPq, column-major
    if constexpr (OPS_LDS) return sPf[idx]; else return ld_off(VfPq, (unsigned)idx);
  };
  const int wave = threadIdx.x / WAVE;
  const int wl = threadIdx.x % WAVE;
  const int tw = wl / NQ;
  const int i = wl - tw * NQ;
  const bool active = tw < TPW;
  const int e = wave * TPW + (active ? tw : 0);
  const int k = blockIdx.x * EPB + e;
  const bool valid = active && k < K;
  const size_t base = (valid ? (size_t)k : 0) * NQ * NFLD;

  // A x (row i) and A^T x (column i) with A = VqPq
  auto mv = [&](double (*in)[NQ], double (&r)[NFLD]) {
    for (int f = 0; f < NFLD; ++f) r[f] = 0.0;
#pragma unroll 8
    for (int j = 0; j < NQ; ++j) {
      const double P = Pv(i + j * NQ);
      for (int f = 0; f < NFLD; ++f) r[f] += P * in[f][j];
    }
  };
  auto mvT = [&](double (*in)[NQ], double (&r)[NFLD]) {
    for (int f = 0; f < NFLD; ++f) r[f] = 0.0;
#pragma unroll 8
    for (int j = 0; j < NQ; ++j) {
      const double P = Pv(j + i * NQ);
      for (int f = 0; f < NFLD; ++f) r[f] += P * in[f][j];
    }
  };

  double U[NFLD], J = 1.0, invJ = 1.0;
  if (valid) {
    J = Jq[(size_t)k * NQ + i];
    invJ = 1.0 / J;
    for (int f = 0; f < NFLD; ++f) U[f] = Q[base + i + f * NQ];
    double V[NFLD];
    entropy_vars<DIM>(U, gamma, V);
    for (int f = 0; f < NFLD; ++f) bufA[e][f][i] = V[f] * J;
  }
  __syncthreads();
  if (valid) {
    double r[NFLD];
    mv(bufA[e], r);
    for (int f = 0; f < NFLD; ++f) bufB[e][f][i] = r[f] * invJ;
  }
  __syncthreads();
  if (valid) {
    double r[NFLD];
    mv(bufB[e], r);
    for (int f = 0; f < NFLD; ++f) bufA[e][f][i] = r[f];  // v~
  }
  __syncthreads();
  // cotangents of v~ through u(v~) and u(VfPq v~): the stage writes rhsQ = u(v~)
  // (v~ already lies in the range of the projection VqPq).  This phase reads
  // only bufA; bufB was last read before the barrier above and bufF is read
  // after the next one, so the results go straight to LDS.
  if (valid) {
    double V[NFLD], Ub[NFLD], g[NFLD];
    for (int f = 0; f < NFLD; ++f) V[f] = bufA[e][f][i];
    for (int f = 0; f < NFLD; ++f) Ub[f] = uqb[base + i + f * NQ];
    conservative_vars_vjp<DIM>(V, gamma, lgm1, Ub, g);
    for (int f = 0; f < NFLD; ++f) bufB[e][f][i] = g[f];
  }
  // face points of the wave's TPW elements, spread over all its lanes
  for (int p = wl; p < TPW * NFN; p += WAVE) {
    const int t = p / NFN;
    const int fn = p - t * NFN;
    const int et = wave * TPW + t;
    const int kt = blockIdx.x * EPB + et;
    if (kt >= K) continue;
    double V[NFLD], Ub[NFLD], g[NFLD];
    for (int f = 0; f < NFLD; ++f) V[f] = 0.0;
    for (int j = 0; j < NQ; ++j) {
      const double P = Pf(fn + j * NFN);
      for (int f = 0; f < NFLD; ++f) V[f] += P * bufA[et][f][j];
    }
    for (int f = 0; f < NFLD; ++f) Ub[f] = ufb[(size_t)kt * NFN * NFLD + fn + f * NFN];
    conservative_vars_vjp<DIM>(V, gamma, lgm1, Ub, g);
    for (int f = 0; f < NFLD; ++f) bufF[et][f][fn] = g[f];
  }
  __syncthreads();
  // v~_bar = g_q + VfPq^T g_f ; then through VqPq diag(1/J) VqPq
  if (valid) {
    double r[NFLD];
    for (int f = 0; f < NFLD; ++f) r[f] = bufB[e][f][i];
    for (int fn = 0; fn < NFN; ++fn) {
      const double P = Pf(fn + i * NFN);
      for (int f = 0; f < NFLD; ++f) r[f] += P * bufF[e][f][fn];
    }
    for (int f = 0; f < NFLD; ++f) bufA[e][f][i] = r[f];
  }
  __syncthreads();
  if (valid) {
    double r[NFLD];
    mvT(bufA[e], r);
    for (int f = 0; f < NFLD; ++f) bufB[e][f][i] = r[f] * invJ;
  }
  __syncthreads();
  double rb[NFLD];
  if (valid) {
    double wb[NFLD], Vb[NFLD], Ub[NFLD];
    mvT(bufB[e], wb);
    for (int f = 0; f < NFLD; ++f) Vb[f] = J * wb[f];
    entropy_vars_vjp<DIM>(U, gamma, Vb, Ub);
    for (int f = 0; f < NFLD; ++f) {
      const size_t id = base + i + f * NQ;
      const double qb = Qb[id] + Ub[f];        // cotangent of Q'
      const double resb_tot = resb[id] + fb * qb;  // Q' = Q + fb res'
      Qb[id] = qb;
      resb[id] = fa * resb_tot;
      rb[f] = fdt * resb_tot;
    }
  }
  __syncthreads();
  // R_bar = -2 VqPq^T diag(1/J) VqPq^T r_bar
  if (valid)
    for (int f = 0; f < NFLD; ++f) bufA[e][f][i] = rb[f];
  __syncthreads();
  if (valid) {
    double r[NFLD];
    mvT(bufA[e], r);
    for (int f = 0; f < NFLD; ++f) bufB[e][f][i] = r[f] * invJ;
  }
  __syncthreads();
  double rbar[NFLD];
  if (valid) {
    double r[NFLD];
    mvT(bufB[e], r);
    for (int f = 0; f < NFLD; ++f) {
      rbar[f] = -2.0 * r[f];
      Rb[base + i + f * NQ] = rbar[f];
    }
  }
  if (!sbar) return;  // uniform: a kernel argument
  // s_bar = VqLq^T R_bar at this element's face points (what sbar_kernel computes)
  __syncthreads();
  if (valid)
    for (int f = 0; f < NFLD; ++f) bufA[e][f][i] = rbar[f];
  __syncthreads();
  for (int p = wl; p < TPW * NFN; p += WAVE) {
    const int t = p / NFN;
    const int fn = p - t * NFN;
    const int et = wave * TPW + t;
    const int kt = blockIdx.x * EPB + et;
    if (kt >= K) continue;
    double v[NFLD];
    for (int f = 0; f < NFLD; ++f) v[f] = 0.0;
#pragma unroll 4
    for (int j = 0; j < NQ; ++j) {
      const double L = ld_off(VqLq, (unsigned)(j + fn * NQ));
      for (int f = 0; f < NFLD; ++f) v[f] += L * bufA[et][f][j];
    }
    for (int f = 0; f < NFLD; ++f) sbar[(size_t)kt * NFN * NFLD + f * NFN + fn] = v[f];
  }
}

// Q_bar += VJP of (u~q, u~f) = (u(VqPq v(Q)), u(VfPq v(Q))) at Q.
template <class S, int TPW>
__global__ void __launch_bounds__(WAVE * 4)
proj_vjp_kernel(int K, const double *__restrict__ VqPq, const double *__restrict__ VfPq,
                const double *__restrict__ Q, const double *__restrict__ uqb,
                const double *__restrict__ ufb, double *__restrict__ Qb, double gamma,
                double lgm1) {
  constexpr int DIM = S::DIM, NQ = S::NQ, NFN = S::NFN, NFLD = S::NFLD;
  constexpr int EPB = 4 * TPW;
  constexpr int MAXF = (NFN + NQ - 1) / NQ;
  __shared__ double bufA[EPB][NFLD][NQ];
  __shared__ double bufB[EPB][NFLD][NQ];
  __shared__ double bufF[EPB][NFLD][NFN];
  const int wave = threadIdx.x / WAVE;
  const int wl = threadIdx.x % WAVE;
  const int tw = wl / NQ;
  const int i = wl - tw * NQ;
  const bool active = tw < TPW;
  const int e = wave * TPW + (active ? tw : 0);
  const int k = blockIdx.x * EPB + e;
  const bool valid = active && k < K;
  const size_t base = (valid ? (size_t)k : 0) * NQ * NFLD;
  const size_t basef = (valid ? (size_t)k : 0) * NFN * NFLD;
  double U[NFLD];
  if (valid) {
    double V[NFLD];
    for (int f = 0; f < NFLD; ++f) U[f] = Q[base + i + f * NQ];
    entropy_vars<DIM>(U, gamma, V);
    for (int f = 0; f < NFLD; ++f) bufA[e][f][i] = V[f];
  }
  __syncthreads();
  double gq[NFLD], gfl[MAXF][NFLD];
  if (valid) {
    double V[NFLD], Ub[NFLD];
    for (int f = 0; f < NFLD; ++f) V[f] = 0.0;
    for (int j = 0; j < NQ; ++j) {
      const double P = ld_off(VqPq, (unsigned)(i + j * NQ));
      for (int f = 0; f < NFLD; ++f) V[f] += P * bufA[e][f][j];
    }
    for (int f = 0; f < NFLD; ++f) Ub[f] = uqb[base + i + f * NQ];
    conservative_vars_vjp<DIM>(V, gamma, lgm1, Ub, gq);
    for (int m = 0; m < MAXF; ++m) {
      const int fn = i + m * NQ;
      if (fn >= NFN) continue;
      for (int f = 0; f < NFLD; ++f) V[f] = 0.0;
      for (int j = 0; j < NQ; ++j) {
        const double P = ld_off(VfPq, (unsigned)(fn + j * NFN));
        for (int f = 0; f < NFLD; ++f) V[f] += P * bufA[e][f][j];
      }
      for (int f = 0; f < NFLD; ++f) Ub[f] = ufb[basef + fn + f * NFN];
      conservative_vars_vjp<DIM>(V, gamma, lgm1, Ub, gfl[m]);
    }
  }
  __syncthreads();
  if (valid) {
    for (int f = 0; f < NFLD; ++f) bufB[e][f][i] = gq[f];
    for (int m = 0; m < MAXF; ++m) {
      const int fn = i + m * NQ;
      if (fn >= NFN) continue;
      for (int f = 0; f < NFLD; ++f) bufF[e][f][fn] = gfl[m][f];
    }
  }
  __syncthreads();
  if (valid) {
    double vb[NFLD], Ub[NFLD];
    for (int f = 0; f < NFLD; ++f) vb[f] = 0.0;
    for (int j = 0; j < NQ; ++j) {
      const double P = ld_off(VqPq, (unsigned)(j + i * NQ));
      for (int f = 0; f < NFLD; ++f) vb[f] += P * bufB[e][f][j];
    }
    for (int fn = 0; fn < NFN; ++fn) {
      const double P = ld_off(VfPq, (unsigned)(fn + i * NFN));
      for (int f = 0; f < NFLD; ++f) vb[f] += P * bufF[e][f][fn];
    }
    entropy_vars_vjp<DIM>(U, gamma, vb, Ub);
    for (int f = 0; f < NFLD; ++f) Qb[base + i + f * NQ] += Ub[f];
  }
}

