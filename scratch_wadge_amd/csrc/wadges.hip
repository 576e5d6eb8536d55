// wadges.hip -- MI355X (gfx950) kernels and C ABI of the WADGES hot path.
//
// Three kernels per LSRK stage, with the roles and argument layouts of the
// reference's OKL kernels (see include/wadges.h for the file:line map):
//
//   vol_kernel     flux-differencing volume term and volume<->face coupling.
//                  FP64 VALU bound.  One team of Nq lanes per element, teams
//                  packed inside one 64-wide wave.  Lane i owns volume row i
//                  in registers; partner rows (volume row j of a symmetric
//                  volume pair, face row f of a volume-face pair) are
//                  accumulated in LDS.  Pairs are visited on a diagonal
//                  schedule (partner = (i + t) mod L) so that at any step the
//                  lanes of a team touch distinct partner rows: each two-point
//                  flux is evaluated once (Nq(Nq+1)/2 + Nq*Nf per element,
//                  versus Nq^2 + 2 Nq Nf in the reference kernel) and the LDS
//                  accumulation order is fixed, so results are deterministic.
//   surf_kernel    entropy-conservative flux + Lax-Friedrichs penalty at the
//                  face points, gathered through mapPq, lifted by VqLq.
//                  HBM bound.
//   update_kernel  weight-adjusted inverse mass, LSRK update, weight-adjusted
//                  entropy projection and u(v) at volume and face points.
//                  HBM / transcendental bound; operators staged in LDS.
//
// Operator tables of vol_kernel are re-laid out per call by a tiny prep kernel
// into the plan's workspace, "diagonal-major": entry [t][i] belongs to the pair
// (i, (i + t) mod L), so at step t the lanes of every team read consecutive
// addresses that all teams of all waves share (L1 hits).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "wadges_device.hpp"
#include "../../include/wadges.h"

namespace wadges {

constexpr int WAVE = 64;

template <int A, int B> struct Max { static constexpr int value = A > B ? A : B; };

// Volume rows per lane of the forward volume block.  With R > 1 a lane holds
// rows l, l + LPE, ..., so a partner record read from LDS, and the partner's
// LDS accumulation, serve R rows, and more elements share a wave: 3D N = 3
// (Nq = 23) packs 5 elements of 12 lanes (60 lanes, 23 of 24 rows used)
// instead of 2 elements of 23 lanes (46 lanes).  R > 1 requires the lanes of an
// element to fit on one face (LPE <= Nfq): the volume-face block then walks
// the faces one at a time.  3D N = 4 (Nq = 44, Nfq = 16) keeps one row per lane:
// three rows (4 elements of 15 lanes) need more registers than the two waves
// per SIMD its LDS allows, and at one wave per SIMD the volume kernel took
// 4.52 ms instead of 3.70 ms at K = 196,608.
constexpr int vol_rows(int dim, int nq, int nfq) { return (dim == 3 && nq == 23 && nfq == 12) ? 2 : 1; }

// p[idx] for a uniform base p and a per-lane index below 2^29: a 32-bit unsigned
// byte offset lets the compiler address through the 64-bit base in scalar
// registers plus a 32-bit lane offset, instead of 64-bit address arithmetic
// per load.
__device__ __forceinline__ double ld_off(const double *p, unsigned idx) {
  return *(const double *)((const char *)p + (idx * 8u));
}

// v, as a value the compiler cannot see through: a loop with this bound keeps
// its compile-time trip count out of the unroller's reach.
__device__ __forceinline__ int opaque_bound(int v) {
  asm volatile("" : "+s"(v));
  return v;
}

// Compile-time shape of one (dim, Nq, Nfq) instance.
template <int DIM_, int NQ_, int NFQ_> struct Shape {
  static constexpr int DIM = DIM_;
  static constexpr int NQ = NQ_;
  static constexpr int NFQ = NFQ_;
  static constexpr int NFACES = DIM + 1;
  static constexpr int NFN = NFQ * NFACES;   // face points per element (Nf)
  static constexpr int NFLD = DIM + 2;
  static constexpr int NG = DIM * DIM;
  static constexpr int L = Max<NQ, NFN>::value;  // diagonal period of the vol-face block
  static constexpr int TVT = NQ / 2 + 1;        // diagonal steps of the vol-vol block
  // node record layouts (structure of arrays in LDS)
  static constexpr int VS = DIM + 5;            // rho, hv[DIM], beta, lrho, lbeta, hq
  static constexpr int VREC = VS + NG;          // + half geofacs
  static constexpr int FREC = VS + DIM;         // + b_d = sum_r hGf_rd nref_r
  // workspace layout (doubles)
  static constexpr int WS_D = 2 * DIM * TVT * NQ;   // D fwd / bwd
  static constexpr int WS_W = 2 * L * NQ;           // +1/2 VqLq, -1/2 VfPq
  // forward volume block: rows per lane, lanes per element, elements per wave
  static constexpr int R = vol_rows(DIM_, NQ_, NFQ_);
  static constexpr int LPE = (NQ + R - 1) / R;
  static constexpr int TPW = WAVE / LPE;
  static constexpr int WS_R = R > 1 ? 2 * NFN * NQ : 0;   // face-walk weights (R > 1)
  // One element per wave with one face point per lane (3D N = 4: Nq = 44, Nf =
  // 64): the volume-face block is walked by face point, every lane busy.
  static constexpr bool COLVF = R == 1 && TPW == 1 && NFN == WAVE && NQ < WAVE;
  static constexpr int WS_C = COLVF ? 2 * NQ * NFN : 0;   // its weights
  static constexpr int WS = WS_D + WS_W + WS_R + WS_C;
  static_assert(R == 1 || (NQ <= NFN && LPE <= NFQ), "R > 1 needs one element's lanes on one face");
};

// ------------------------------------------------------------------ pair schedules
//
// Volume-volume block: lane i visits j = (i + t) mod NQ, t = 0..NQ/2 (for even NQ
// only lanes i < NQ/2 take t = NQ/2), so every unordered pair and the diagonal
// is evaluated once, and at fixed t the partners j are distinct.
//
// Volume-face block, "face-chunked" when NQ <= NFN: lanes form groups of NFQ
// (g = i / NFQ, o = i % NFQ); step t = c*NFQ + m visits face (c + g) mod NFACES,
// point (o + m) mod NFQ.  Within a step the partners are distinct (distinct
// faces across groups, distinct points within a group) and within a chunk c a
// lane's face is fixed, so its own-side normal term is computed once per face.
// Otherwise (2D N >= 4, where NQ > NFN) the plain diagonal fn = (i + t) mod L.
template <class S> __host__ __device__ constexpr bool face_chunked() { return S::NQ <= S::NFN; }

template <class S> __device__ __forceinline__ int vf_partner(int t, int i) {
  if constexpr (face_chunked<S>()) {
    const int c = t / S::NFQ, m = t - (t / S::NFQ) * S::NFQ;
    const int g = i / S::NFQ, o = i - (i / S::NFQ) * S::NFQ;
    int face = c + g;
    if (face >= S::NFACES) face -= S::NFACES;
    int node = o + m;
    if (node >= S::NFQ) node -= S::NFQ;
    return face * S::NFQ + node;
  } else {
    int f = i + t;
    if (f >= S::L) f -= S::L;
    return f;  // >= NFN means "no pair at this step"
  }
}

template <class S> constexpr int vf_steps() { return face_chunked<S>() ? S::NFN : S::L; }

// The volume-face block takes the face records in parts, holding only the
// current part's records in LDS.  With one row per lane, face-chunked
// tetrahedra whose lanes form at most two groups visit the faces in two halves
// (faces 0-1, then 2-3), the two groups on different faces of the half at every
// step (the steps t and their partners are those of vf_partner, taken in another
// order).  Face-chunked triangles whose lanes form two groups (N = 1, 2) walk
// the faces one at a time, both groups on the same face (their face-accumulator
// additions then go in two instructions, one per group): the smaller LDS
// footprint matters to the ensemble kernel, whose LDS per wave is that of its
// largest shape.  With R > 1 rows per lane an element's lanes form one group,
// which walks the faces one at a time.
template <class S> constexpr int face_groups() { return (S::NQ + S::NFQ - 1) / S::NFQ; }
template <class S> constexpr int face_parts() {
  if constexpr (S::R > 1) return S::NFACES;
  if constexpr (face_chunked<S>() && S::DIM == 2 && face_groups<S>() == 2) return S::NFACES;
  return (face_chunked<S>() && S::NFACES == 4 && face_groups<S>() <= 2) ? 2 : 1;
}

// ------------------------------------------------------------------ operator prep

template <class S>
__global__ void prep_vol_tables(const double *__restrict__ Drq, const double *__restrict__ Dsq,
                                const double *__restrict__ Dtq, const double *__restrict__ VqLq,
                                const double *__restrict__ VfPq, double *__restrict__ ws) {
  constexpr int NQ = S::NQ, NFN = S::NFN, TVT = S::TVT, DIM = S::DIM, T = vf_steps<S>();
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < S::WS_D) {
    const int i = idx % NQ;
    const int t = (idx / NQ) % TVT;
    const int r = (idx / (NQ * TVT)) % DIM;
    const int s = idx / (NQ * TVT * DIM);
    int j = i + t;
    if (j >= NQ) j -= NQ;
    const double *D = r == 0 ? Drq : (r == 1 ? Dsq : Dtq);
    ws[idx] = s == 0 ? D[i + j * NQ] : D[j + i * NQ];
  } else if (idx < S::WS_D + 2 * T * NQ) {
    const int k = idx - S::WS_D;
    const int i = k % NQ;
    const int t = (k / NQ) % T;
    const int s = k / (NQ * T);
    const int f = vf_partner<S>(t, i);
    double v = 0.0;
    if (f < NFN) v = s == 0 ? 0.5 * VqLq[i + f * NQ] : -0.5 * VfPq[f + i * NFN];
    ws[idx] = v;
  } else if (S::R > 1 && idx >= S::WS_D + S::WS_W && idx < S::WS) {
    // face walk of R > 1: step t = face * NFQ + m, row i on lane l = i mod LPE
    // visits point (l + m) mod NFQ of that face
    const int k = idx - S::WS_D - S::WS_W;
    const int i = k % NQ;
    const int t = (k / NQ) % NFN;
    const int s = k / (NQ * NFN);
    int q = i % S::LPE + t % S::NFQ;
    if (q >= S::NFQ) q -= S::NFQ;
    const int f = (t / S::NFQ) * S::NFQ + q;
    ws[idx] = s == 0 ? 0.5 * VqLq[i + f * NQ] : -0.5 * VfPq[f + i * NFN];
  } else if (S::COLVF && idx >= S::WS_D + S::WS_W + S::WS_R && idx < S::WS) {
    // face-point walk: step t, lane L (face point L) visits row (L mod NQ + t) mod NQ
    const int k = idx - S::WS_D - S::WS_W - S::WS_R;
    const int L = k % NFN;
    const int t = (k / NFN) % NQ;
    const int s = k / (NFN * NQ);
    int i = L % NQ + t;
    if (i >= NQ) i -= NQ;
    ws[idx] = s == 0 ? 0.5 * VqLq[i + L * NQ] : -0.5 * VfPq[L + i * NFN];
  }
}

// ------------------------------------------------------------------ volume kernel

// LDS per element: the volume accumulator, then one region used first for the
// volume-node records (volume-volume block) and then for the face-node records
// and face accumulator (volume-face block).
template <class S> struct VolSmem {
  static constexpr int H = face_parts<S>();
  static constexpr int NH = S::NFN / H;                    // face records held at a time
  static constexpr int acc_v = S::NFLD * S::NQ;
  static constexpr int phase1 = S::VREC * S::NQ;
  static constexpr int phase2 = S::FREC * NH + S::NFLD * S::NFN;
  static constexpr int per_elem = acc_v + (phase1 > phase2 ? phase1 : phase2);
  static constexpr int accF_off = acc_v + S::FREC * NH;    // face accumulators [NFLD][NFN]
  static_assert(per_elem >= 2 * acc_v, "the fused stage's update needs two [NFLD][NQ] buffers");
};

// Waves per SIMD that the LDS of a one-wave workgroup of `bytes` allows (160 KiB
// of LDS per CU, four SIMDs), rounded up.  Kernels whose occupancy LDS limits
// declare it, so that the compiler may spend the registers that occupancy
// leaves free on latency hiding (batched loads, unrolled row products) instead
// of aiming at an occupancy the LDS cannot reach.
constexpr int lds_waves_per_simd(int bytes) {
  const int wg = (160 * 1024) / (bytes > 0 ? bytes : 1);
  const int w = (wg + 3) / 4;
  return w < 1 ? 1 : (w > 8 ? 8 : w);
}

template <class S> constexpr int vol_waves() {
  return lds_waves_per_simd(S::TPW * VolSmem<S>::per_elem * 8 + S::NFACES * S::DIM * 8);
}
// Lower bound handed to the compiler: at most four waves (128 VGPRs), so that
// shapes whose LDS would allow more are not squeezed into spills.
template <class S> constexpr int vol_waves_min() { return vol_waves<S>() < 4 ? vol_waves<S>() : 4; }

template <int DIM>
__device__ __forceinline__ Node<DIM> load_node(const double *rec, int stride, int j) {
  Node<DIM> n;
  n.rho = rec[0 * stride + j];
  for (int d = 0; d < DIM; ++d) n.hv[d] = rec[(1 + d) * stride + j];
  n.beta = rec[(DIM + 1) * stride + j];
  n.lrho = rec[(DIM + 2) * stride + j];
  n.lbeta = rec[(DIM + 3) * stride + j];
  n.hq = rec[(DIM + 4) * stride + j];
  return n;
}

template <int DIM>
__device__ __forceinline__ void store_node(double *rec, int stride, int j, const Node<DIM> &n) {
  rec[0 * stride + j] = n.rho;
  for (int d = 0; d < DIM; ++d) rec[(1 + d) * stride + j] = n.hv[d];
  rec[(DIM + 1) * stride + j] = n.beta;
  rec[(DIM + 2) * stride + j] = n.lrho;
  rec[(DIM + 3) * stride + j] = n.lbeta;
  rec[(DIM + 4) * stride + j] = n.hq;
}

// Where the other teams of the wave find their data (for phases that spread the
// face points of all elements of a wave over all 64 lanes).
struct WaveCtx {
  int lane, kfirst, K;
  const double *Uf;    // face state of element 0 of the mesh
  const double *vfg;   // face geofacs of element 0
  double *smem;        // the workgroup's dynamic LDS (team t at t * per_elem)
};

// The volume term of one element (both coupling blocks).  Lane l of the
// element's team owns rows l + r * LPE (r < R, those below NQ).  On return
// acc[r] holds the register part of its row r, accV the LDS part (partner
// contributions), accF the face rows (rhsQf).  Uq is the element's projected
// volume state, vg its volume geofacs; the face state and geofacs of the wave's
// elements come through wc.  Every lane of the workgroup must call this
// (barriers).
template <class S>
__device__ __forceinline__ void vol_body(bool valid, int l, const double *Uq, const double *vg,
                                         const double (*s_nref)[S::DIM], const double *ws,
                                         double *accV, double gm1, double inv2gm1,
                                         double (&acc)[S::R][S::NFLD], const WaveCtx &wc) {
  constexpr int DIM = S::DIM, NQ = S::NQ, NFQ = S::NFQ, NFN = S::NFN, NFLD = S::NFLD;
  constexpr int NG = S::NG, TVT = S::TVT, VS = S::VS, R = S::R, LPE = S::LPE;
  constexpr int T = vf_steps<S>();
  constexpr int H = VolSmem<S>::H, NH = VolSmem<S>::NH;
  double *vrec = accV + NFLD * NQ;         // phase 1
  double *frec = vrec;                      // phase 2: records of NH face points
  double *accF = accV + VolSmem<S>::accF_off;

  bool rv[R];                               // row r of this lane exists
  for (int r = 0; r < R; ++r) rv[r] = valid && (r == 0 || l + r * LPE < NQ);

  // Smooth waves (3D): every node of the wave's elements within a band of
  // element 0's node 0 (read from LDS) in rho and beta; the loops then skip the
  // per-pair branch test of the log means, and the face records take their
  // logarithms only when a wave is not smooth (same bits).  Taken where it
  // pays: with two rows per lane (3D N = 3: 1.6 % and 2.2 % faster on the
  // Taylor-Green state); at 3D N = 4 the test measured 1.5 % slower.
  constexpr bool SMOOTH_LOOPS = DIM == 3 && R > 1;
  constexpr bool LAZY = SMOOTH_LOOPS;   // face-record logarithms only where needed

  // ---------------- phase 1: volume records, volume-volume block
  Node<DIM> me[R];
  double hG[R][NG];
  for (int r = 0; r < R; ++r) {
    if (!rv[r]) continue;
    const int i = l + r * LPE;
    double U[NFLD];
    for (int f = 0; f < NFLD; ++f) U[f] = Uq[i + f * NQ];
    me[r] = make_node<DIM>(U, gm1);
    for (int c = 0; c < NG; ++c) hG[r][c] = 0.5 * vg[i + c * NQ];
    store_node<DIM>(vrec, NQ, i, me[r]);
    for (int c = 0; c < NG; ++c) vrec[(VS + c) * NQ + i] = hG[r][c];
    for (int f = 0; f < NFLD; ++f) accV[f * NQ + i] = 0.0;
  }
  __syncthreads();

  const double rho_ref = wc.smem[NFLD * NQ];
  const double beta_ref = wc.smem[NFLD * NQ + (DIM + 1) * NQ];
  bool smooth_v = false;
  if constexpr (SMOOTH_LOOPS) {
    bool ok = true;
    for (int r = 0; r < R; ++r)
      if (rv[r]) ok = ok && in_smooth_band(me[r].rho, me[r].beta, rho_ref, beta_ref);
    smooth_v = uniform_flag(wave_all(ok));
  }

  for (int r = 0; r < R; ++r)
    for (int f = 0; f < NFLD; ++f) acc[r][f] = 0.0;

  // 3D N = 2 (Nq = 11): its short loops (6 steps here and in the face walk) are
  // kept rolled; unrolled completely they spill even at 168 registers.
  constexpr bool ROLLED = S::DIM == 3 && NQ == 11;
  int vv_steps = TVT;
  if constexpr (ROLLED) vv_steps = opaque_bound(TVT);
  if (valid) {
    // With R > 1 the operator entries of step t + 1 are loaded while step t
    // computes (one row per lane has no registers to spare for it).  (Two
    // buffers used in turn, as in the face walk below, spilled here.)
    constexpr bool PIPE = R > 1 || DIM == 2;
    const double *Dfw = ws;
    const double *Dbw = ws + DIM * TVT * NQ;
    int iw[R];                        // table row (row 0's for rows that do not exist)
    for (int r = 0; r < R; ++r) iw[r] = rv[r] ? l + r * LPE : 0;
    auto fetch = [&](int t, double (&Df)[R][DIM], double (&Db)[R][DIM]) {
      for (int r = 0; r < R; ++r)
        for (int q = 0; q < DIM; ++q) {
          Df[r][q] = ld_off(Dfw, (unsigned)((q * TVT + t) * NQ + iw[r]));
          Db[r][q] = ld_off(Dbw, (unsigned)((q * TVT + t) * NQ + iw[r]));
        }
    };
    auto vv_step = [&](int t, const double (&Df)[R][DIM], const double (&Db)[R][DIM]) {
      for (int r = 0; r < R; ++r) {
        if (!rv[r]) continue;
        const int i = l + r * LPE;
        if (2 * t == NQ && i >= NQ / 2) continue;  // even NQ: middle pair once
        int j = i + t;
        if (j >= NQ) j -= NQ;
        const Node<DIM> pj = load_node<DIM>(vrec, NQ, j);
        double Ga[NG];
        for (int c = 0; c < NG; ++c) Ga[c] = hG[r][c] + vrec[(VS + c) * NQ + j];
        const PairFlux<DIM> P = pair_flux<DIM, false, (R > 1)>(me[r], pj, inv2gm1, smooth_v);
        double Dx[DIM];
        for (int d = 0; d < DIM; ++d) Dx[d] = 0.0;
        for (int q = 0; q < DIM; ++q)
          for (int d = 0; d < DIM; ++d) Dx[d] = Dx[d] + Df[r][q] * Ga[q * DIM + d];
        double c[NFLD];
        flux_dot<DIM>(P, Dx, c);
        for (int f = 0; f < NFLD; ++f) acc[r][f] += c[f];
        if (t > 0) {
          for (int d = 0; d < DIM; ++d) Dx[d] = 0.0;
          for (int q = 0; q < DIM; ++q)
            for (int d = 0; d < DIM; ++d) Dx[d] = Dx[d] + Db[r][q] * Ga[q * DIM + d];
          flux_dot<DIM>(P, Dx, c);
          for (int f = 0; f < NFLD; ++f) lds_add(&accV[f * NQ + j], c[f]);
        }
      }
    };
    {
      if constexpr (PIPE) {
        double nDf[R][DIM], nDb[R][DIM];
        fetch(0, nDf, nDb);
        for (int t = 0; t < TVT; ++t) {
          if (2 * t > NQ) break;
          double Df[R][DIM], Db[R][DIM];
          for (int r = 0; r < R; ++r)
            for (int q = 0; q < DIM; ++q) {
              Df[r][q] = nDf[r][q];
              Db[r][q] = nDb[r][q];
            }
          if (t + 1 < TVT && 2 * (t + 1) <= NQ) fetch(t + 1, nDf, nDb);
          vv_step(t, Df, Db);
        }
      } else {
        for (int t = 0; t < vv_steps; ++t) {
          if (2 * t > NQ) break;
          double Df[R][DIM], Db[R][DIM];
          fetch(t, Df, Db);
          vv_step(t, Df, Db);
        }
      }
    }
  }
  __syncthreads();

  // ---------------- phase 2: face records (overwrite the volume records),
  // volume-face block, in H parts of NH face points.  Records and accumulator
  // clearing are spread over the whole wave (all teams' face points together).
  constexpr int TPW = S::TPW, PER = VolSmem<S>::per_elem;
  // own-side part of the averaged normal of row r on a face: sum_q (G_qd / 2) nref_q
  auto own_normal = [&](int r, int face, double (&a)[DIM]) {
    for (int d = 0; d < DIM; ++d) {
      double v = 0.0;
      for (int q = 0; q < DIM; ++q) v = v + hG[r][q * DIM + d] * s_nref[face][q];
      a[d] = v;
    }
  };
  if constexpr (S::COLVF) {
    // Lane L takes face point L of the wave's one element and walks the rows:
    // at step t row i = (L mod NQ + t) mod NQ.  The face point's record and
    // its accumulator stay in registers; the volume records stay in LDS, with
    // each row's own-side normal part for the four faces written from the half
    // geofacs on (the face accumulator, written behind them, is stored only
    // after the walk).  Row contributions go to the LDS accumulator, lanes
    // L < NQ and L >= NQ in two instructions (their rows coincide), so no two
    // lanes of an instruction add into one address and the order of the
    // additions is fixed.  The row's register part of the volume-volume block
    // joins its LDS part now (registers for the walk; the row sum is then the
    // LDS value alone).  A step's loads are not issued a step ahead: that
    // measured slower here than the compiler's schedule of the plain loop.
    static_assert(NFLD * NQ + (VS + S::NFACES * DIM) * NQ <= PER, "normal parts inside the element's LDS");
    if (rv[0]) {
      for (int face = 0; face < S::NFACES; ++face) {
        double a[DIM];
        own_normal(0, face, a);
        for (int d = 0; d < DIM; ++d) vrec[(VS + face * DIM + d) * NQ + l] = a[d];
      }
      for (int f = 0; f < NFLD; ++f) {
        accV[f * NQ + l] = acc[0][f] + accV[f * NQ + l];
        acc[0][f] = 0.0;
      }
    }
    const int L = wc.lane;
    const bool cv = wc.kfirst < wc.K;
    const int face = L / NFQ;
    Node<DIM> nf;
    double bf[DIM];
    if (cv) {
      const double *Uft = wc.Uf + (size_t)wc.kfirst * NFN * NFLD;
      const double *vfgt = wc.vfg + (size_t)wc.kfirst * NFN * NG;
      double U[NFLD];
      for (int f = 0; f < NFLD; ++f) U[f] = Uft[L + f * NFN];
      nf = make_node<DIM>(U, gm1);
      for (int d = 0; d < DIM; ++d) {
        double b = 0.0;
        for (int q = 0; q < DIM; ++q) b = b + 0.5 * vfgt[L + (q * DIM + d) * NFN] * s_nref[face][q];
        bf[d] = b;
      }
    }
    __syncthreads();
    double sF[NFLD];
    for (int f = 0; f < NFLD; ++f) sF[f] = 0.0;
    if (cv) {
      const double *Wc1 = ws + S::WS_D + S::WS_W + S::WS_R;
      const double *Wc2 = Wc1 + NQ * NFN;
      const int i0 = L < NQ ? L : L - NQ;
      const bool second = L >= NQ;
      struct ColStep {
        Node<DIM> pv;
        double a[DIM], w1, w2;
      };
      auto fetch = [&](int t, ColStep &x) {
        int i = i0 + t;
        if (i >= NQ) i -= NQ;
        x.pv = load_node<DIM>(vrec, NQ, i);
        for (int d = 0; d < DIM; ++d) x.a[d] = vrec[(VS + face * DIM + d) * NQ + i];
        x.w1 = ld_off(Wc1, (unsigned)(t * NFN + L));
        x.w2 = ld_off(Wc2, (unsigned)(t * NFN + L));
      };
      auto step = [&](int t, const ColStep &x) {
        int i = i0 + t;
        if (i >= NQ) i -= NQ;
        double n[DIM];
        for (int d = 0; d < DIM; ++d) n[d] = x.a[d] + bf[d];
        const PairFlux<DIM> P = pair_flux<DIM>(x.pv, nf, inv2gm1);
        double Fn[NFLD];
        flux_dot<DIM>(P, n, Fn);
        double c[NFLD];
        for (int f = 0; f < NFLD; ++f) {
          sF[f] += x.w2 * Fn[f];
          c[f] = x.w1 * Fn[f];
        }
        if (!second)
          for (int f = 0; f < NFLD; ++f) lds_add(&accV[f * NQ + i], c[f]);
        if (second)
          for (int f = 0; f < NFLD; ++f) lds_add_second(&accV[f * NQ + i], c[f]);
      };
      for (int t = 0; t < NQ; ++t) {
        ColStep x;
        fetch(t, x);
        step(t, x);
      }
    }
    __syncthreads();
    if (cv)
      for (int f = 0; f < NFLD; ++f) accF[f * NFN + L] = sF[f];
    __syncthreads();
    return;
  }
  for (int p = wc.lane; p < TPW * NFN * NFLD; p += WAVE) {
    const int t = p / (NFN * NFLD);
    wc.smem[t * PER + VolSmem<S>::accF_off + (p - t * NFN * NFLD)] = 0.0;
  }
  for (int h = 0; h < H; ++h) {
    bool ok_f = true;
    for (int p = wc.lane; p < TPW * NH; p += WAVE) {
      const int t = p / NH;
      const int fl = p - t * NH;
      const int kt = wc.kfirst + t;
      if (kt >= wc.K) continue;
      const double *Uft = wc.Uf + (size_t)kt * NFN * NFLD;
      const double *vfgt = wc.vfg + (size_t)kt * NFN * NG;
      double *frect = wc.smem + t * PER + NFLD * NQ;
      const int fn = h * NH + fl;
      double U[NFLD];
      for (int f = 0; f < NFLD; ++f) U[f] = Uft[fn + f * NFN];
      const Node<DIM> nf = make_node<DIM, !LAZY>(U, gm1);
      if constexpr (SMOOTH_LOOPS) ok_f = ok_f && in_smooth_band(nf.rho, nf.beta, rho_ref, beta_ref);
      store_node<DIM>(frect, NH, fl, nf);
      const int face = fn / NFQ;
      for (int d = 0; d < DIM; ++d) {
        double b = 0.0;
        for (int q = 0; q < DIM; ++q) b = b + 0.5 * vfgt[fn + (q * DIM + d) * NFN] * s_nref[face][q];
        frect[(VS + d) * NH + fl] = b;
      }
    }
    __syncthreads();
    const bool smooth_f = SMOOTH_LOOPS && uniform_flag(smooth_v && wave_all(ok_f));
    if (LAZY && !smooth_f) {
      // the log branch may be taken: logarithms of this part's face records
      for (int p = wc.lane; p < TPW * NH; p += WAVE) {
        const int t = p / NH;
        const int fl = p - t * NH;
        if (wc.kfirst + t >= wc.K) continue;
        double *frect = wc.smem + t * PER + NFLD * NQ;
        frect[(DIM + 2) * NH + fl] = log(frect[0 * NH + fl]);
        frect[(DIM + 3) * NH + fl] = log(frect[(DIM + 1) * NH + fl]);
      }
      __syncthreads();
    }
    if (valid) {
      if constexpr (R > 1) {
        // part h is face h; at step m lane l visits its point (l + m) mod NFQ,
        // and the one record read serves all rows of the lane.  The partner
        // record and the weights of step m + 1 are loaded while step m computes.
        const double *W1 = ws + S::WS_D + S::WS_W;
        const double *W2 = W1 + NFN * NQ;
        double a[R][DIM];
        int iw[R];                        // weight row (row 0's for rows that do not exist)
        for (int r = 0; r < R; ++r) {
          iw[r] = rv[r] ? l + r * LPE : 0;
          if (rv[r]) own_normal(r, h, a[r]);
        }
        // partner record and weights of one step, in two buffers used in turn
        // (steps taken in pairs: NFQ is even), so no value is copied between steps
        struct VfStep {
          Node<DIM> pf;
          double b[DIM], w1[R], w2[R];
        };
        static_assert(NFQ % 2 == 0, "face-walk steps taken in pairs");
        auto fetch = [&](int m, VfStep &x) {
          int q = l + m;
          if (q >= NFQ) q -= NFQ;
          x.pf = load_node<DIM>(frec, NH, q);
          for (int d = 0; d < DIM; ++d) x.b[d] = frec[(VS + d) * NH + q];
          const int t = h * NFQ + m;
          for (int r = 0; r < R; ++r) {
            x.w1[r] = ld_off(W1, (unsigned)(t * NQ + iw[r]));
            x.w2[r] = ld_off(W2, (unsigned)(t * NQ + iw[r]));
          }
        };
        auto vf_step = [&](int m, const VfStep &x) {
          int q = l + m;
          if (q >= NFQ) q -= NFQ;
          double sF[NFLD];
          for (int f = 0; f < NFLD; ++f) sF[f] = 0.0;
          for (int r = 0; r < R; ++r) {
            if (!rv[r]) continue;
            double n[DIM];
            for (int d = 0; d < DIM; ++d) n[d] = a[r][d] + x.b[d];
            const PairFlux<DIM> P = pair_flux<DIM, false, (R > 1)>(me[r], x.pf, inv2gm1, smooth_f);
            double Fn[NFLD];
            flux_dot<DIM>(P, n, Fn);
            for (int f = 0; f < NFLD; ++f) {
              acc[r][f] += x.w1[r] * Fn[f];
              sF[f] += x.w2[r] * Fn[f];
            }
          }
          for (int f = 0; f < NFLD; ++f) lds_add(&accF[f * NFN + h * NFQ + q], sF[f]);
        };
        VfStep s0, s1;
        fetch(0, s0);
        for (int m = 0; m < NFQ; m += 2) {
          fetch(m + 1, s1);
          vf_step(m, s0);
          if (m + 2 < NFQ) fetch(m + 2, s0);
          vf_step(m + 1, s1);
        }
      } else {
        const int i = l;
        const double *W1 = ws + S::WS_D;
        const double *W2 = W1 + T * NQ;
        // one volume-face pair: step t, partner face point fn (record fl), own-side normal part a
        // groups of lanes that meet on one face point in a step (faces per part
        // fewer than groups) add into the face accumulator in two instructions
        constexpr bool SPLIT = face_chunked<S>() && face_groups<S>() > S::NFACES / H;
        static_assert(!SPLIT || face_groups<S>() == 2, "two lane groups at most per face");
        auto vf_pair = [&](int fn, int fl, const double (&a)[DIM], bool second, double w1,
                           double w2) {
          const Node<DIM> pf = load_node<DIM>(frec, NH, fl);
          double n[DIM];
          for (int d = 0; d < DIM; ++d) n[d] = a[d] + frec[(VS + d) * NH + fl];
          const PairFlux<DIM> P = pair_flux<DIM>(me[0], pf, inv2gm1, smooth_f);
          double Fn[NFLD];
          flux_dot<DIM>(P, n, Fn);
          for (int f = 0; f < NFLD; ++f) acc[0][f] += w1 * Fn[f];
          if (!SPLIT || !second)
            for (int f = 0; f < NFLD; ++f) lds_add(&accF[f * NFN + fn], w2 * Fn[f]);
          if (SPLIT && second)
            for (int f = 0; f < NFLD; ++f) lds_add_second(&accF[f * NFN + fn], w2 * Fn[f]);
        };
        if constexpr (face_chunked<S>()) {
          const int g = i / NFQ, o = i - (i / NFQ) * NFQ;
          constexpr int FH = S::NFACES / H;            // faces per part
          int fh_steps = FH, m_steps = NFQ;
          if constexpr (ROLLED) {
            fh_steps = opaque_bound(FH);
            m_steps = opaque_bound(NFQ);
          }
          {
            for (int u = 0; u < fh_steps; ++u) {
              // face of this group at part-step u; the chunk c of vf_partner that visits it
              int fh = u + g;
              while (fh >= FH) fh -= FH;
              const int face = h * FH + fh;
              int c = face - g;
              if (c < 0) c += S::NFACES;
              double a[DIM];
              own_normal(0, face, a);
              // 2D: a step's operator entries are loaded during the step before
              constexpr bool VFPIPE = DIM == 2;
              double w1n = 0.0, w2n = 0.0;
              if (VFPIPE) {
                w1n = ld_off(W1, (unsigned)(c * NFQ * NQ + i));
                w2n = ld_off(W2, (unsigned)(c * NFQ * NQ + i));
              }
              for (int m = 0; m < m_steps; ++m) {
                int node = o + m;
                if (node >= NFQ) node -= NFQ;
                const int fn = face * NFQ + node;
                const int t = c * NFQ + m;
                double w1, w2;
                if (VFPIPE) {
                  w1 = w1n;
                  w2 = w2n;
                  if (m + 1 < m_steps) {
                    w1n = ld_off(W1, (unsigned)((t + 1) * NQ + i));
                    w2n = ld_off(W2, (unsigned)((t + 1) * NQ + i));
                  }
                } else {
                  w1 = ld_off(W1, (unsigned)(t * NQ + i));
                  w2 = ld_off(W2, (unsigned)(t * NQ + i));
                }
                vf_pair(fn, fn - h * NH, a, g > 0, w1, w2);
              }
            }
          }
        } else {
          for (int t = 0; t < T; ++t) {
            const int fn = vf_partner<S>(t, i);
            if (fn >= NFN) continue;
            double a[DIM];
            own_normal(0, fn / NFQ, a);
            vf_pair(fn, fn, a, false, ld_off(W1, (unsigned)(t * NQ + i)),
                    ld_off(W2, (unsigned)(t * NQ + i)));
          }
        }
      }
    }
    if (H > 1) __syncthreads();   // records are overwritten by the next part
  }
  __syncthreads();
}

template <class S>
__device__ __forceinline__ void load_nref(int lane, const double *nrJ, const double *nsJ,
                                          const double *ntJ, double (*s_nref)[S::DIM]) {
  if (lane < S::NFACES) {
    s_nref[lane][0] = nrJ[lane * S::NFQ];
    s_nref[lane][1] = nsJ[lane * S::NFQ];
    if constexpr (S::DIM == 3) s_nref[lane][2] = ntJ[lane * S::NFQ];
  }
}

template <class S>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(vol_waves_min<S>(), vol_waves<S>())))
vol_kernel(int K, const double *__restrict__ vgeo, const double *__restrict__ vfgeo,
           const double *__restrict__ nrJ, const double *__restrict__ nsJ,
           const double *__restrict__ ntJ, const double *__restrict__ ws,
           const double *__restrict__ Qf, double *rhsQ, double *__restrict__ rhsQf,
           double gm1, double inv2gm1) {
  constexpr int NQ = S::NQ, NFN = S::NFN, NFLD = S::NFLD, NG = S::NG;
  constexpr int R = S::R, LPE = S::LPE, TPW = S::TPW;
  constexpr int PER = VolSmem<S>::per_elem;
  extern __shared__ double smem[];
  __shared__ double s_nref[S::NFACES][S::DIM];

  const int lane = threadIdx.x;
  const int team = lane / LPE;
  const int l = lane - team * LPE;
  const bool active = team < TPW;
  const int k = blockIdx.x * TPW + team;
  const bool valid = active && k < K;
  const size_t kk = valid ? (size_t)k : 0;

  load_nref<S>(lane, nrJ, nsJ, ntJ, s_nref);
  double *accV = smem + (active ? team : 0) * PER;
  double acc[R][NFLD];
  const WaveCtx wc{lane, (int)blockIdx.x * TPW, K, Qf, vfgeo, smem};
  vol_body<S>(valid, l, rhsQ + kk * NQ * NFLD, vgeo + kk * NQ * NG, s_nref, ws, accV, gm1,
              inv2gm1, acc, wc);
  const double *accF = accV + VolSmem<S>::accF_off;

  if (valid) {
    double *dst = rhsQ + (size_t)k * NQ * NFLD;
    for (int r = 0; r < R; ++r) {
      const int i = l + r * LPE;
      if (r > 0 && i >= NQ) continue;
      for (int f = 0; f < NFLD; ++f) dst[i + f * NQ] = acc[r][f] + accV[f * NQ + i];
    }
    double *dstf = rhsQf + (size_t)k * NFN * NFLD;
    for (int fn = l; fn < NFN; fn += LPE)
      for (int f = 0; f < NFLD; ++f) dstf[fn + f * NFN] = accF[f * NFN + fn];
  }
}

// ------------------------------------------------------------------ surface kernel

// Elements per surf_kernel workgroup: the face points (then the lift rows) of
// all of them are spread over the workgroup's 256 lanes, so no lane idles for
// the shape of one element (3D N = 3: 48 face points and 23 rows per element
// left 16 and 41 lanes of a wave idle).  256 / gcd(256, Nf) makes the face
// points fill whole passes; at least 16, at most 64 (LDS: 41 KB at 3D N = 4).
constexpr int surf_epb(int nfn) {
  int a = 256, b = nfn;
  while (b) { const int t = a % b; a = b; b = t; }
  const int e = 256 / a;
  return e < 16 ? 16 : (e > 64 ? 64 : e);
}

template <class S, int EPB>
__global__ void __launch_bounds__(WAVE * 4)
surf_kernel(int K, const double *__restrict__ fgeo, const uint32_t *__restrict__ mapPq,
            const double *__restrict__ VqLq, const double *__restrict__ Qf,
            const double *__restrict__ rhsQf, double *__restrict__ rhsQ, double gamma,
            double gm1, double inv2gm1, double ltau) {
  constexpr int DIM = S::DIM, NQ = S::NQ, NFN = S::NFN, NFLD = S::NFLD, NFG = DIM + 1;
  constexpr int NT = WAVE * 4;
  __shared__ double sfS[EPB][NFLD][NFN];
  const int k0 = blockIdx.x * EPB;
  const int ne = (K - k0 < EPB) ? K - k0 : EPB;   // elements of this workgroup, >= 1

  for (int p = threadIdx.x; p < ne * NFN; p += NT) {
    const int e = p / NFN;
    const int fn = p - e * NFN;
    const int k = k0 + e;
    const size_t idM = (size_t)fn + (size_t)k * NFN * NFLD;
    const size_t idP = mapPq[(size_t)k * NFN + fn];
    double UM[NFLD], UP[NFLD];
    for (int f = 0; f < NFLD; ++f) {
      UM[f] = Qf[idM + (size_t)f * NFN];
      UP[f] = Qf[idP + (size_t)f * NFN];
    }
    const double *g = fgeo + (size_t)k * NFN * NFG;
    double nJ[DIM];
    for (int d = 0; d < DIM; ++d) nJ[d] = g[fn + d * NFN];
    const double sJ = g[fn + DIM * NFN];
    // 3D: the logarithms are taken only if a lane of the wave needs the log branch
    const Node<DIM> a = make_node<DIM, DIM == 2>(UM, gm1);
    const Node<DIM> b = make_node<DIM, DIM == 2>(UP, gm1);
    // wave speed |u| + sqrt(gamma p / rho)  (okl/Euler3D.okl:3316-3321); 2 hq = |u|^2 exactly
    const double pM = gm1 * (UM[NFLD - 1] - 0.5 * UM[0] * (2.0 * a.hq));
    const double pP = gm1 * (UP[NFLD - 1] - 0.5 * UP[0] * (2.0 * b.hq));
    const double lamM = sqrt(2.0 * a.hq) + sqrt(gamma * pM / UM[0]);
    const double lamP = sqrt(2.0 * b.hq) + sqrt(gamma * pP / UP[0]);
    const double Lfc = ltau * fmax(lamM, lamP) * sJ;
    const PairFlux<DIM> P = pair_flux<DIM, DIM == 3, (S::R > 1)>(a, b, inv2gm1);
    double Fn[NFLD];
    flux_dot<DIM>(P, nJ, Fn);
    const double *rf = rhsQf + (size_t)k * NFN * NFLD;
    for (int f = 0; f < NFLD; ++f)
      sfS[e][f][fn] = 0.5 * Fn[f] + rf[fn + f * NFN] - Lfc * (UP[f] - UM[f]);
  }
  __syncthreads();
  for (int p = threadIdx.x; p < ne * NQ; p += NT) {
    const int e = p / NQ;
    const int row = p - e * NQ;
    const int k = k0 + e;
    double val[NFLD];
    for (int f = 0; f < NFLD; ++f) val[f] = 0.0;
#pragma unroll 2
    for (int j = 0; j < NFN; ++j) {
      const double Lij = ld_off(VqLq, (unsigned)(row + j * NQ));
      for (int f = 0; f < NFLD; ++f) val[f] += Lij * sfS[e][f][j];
    }
    double *dst = rhsQ + (size_t)k * NQ * NFLD;
    for (int f = 0; f < NFLD; ++f) dst[row + f * NQ] += val[f];
  }
}

// ------------------------------------------------------------------ FP64 matrix-core tiles
//
// v_mfma_f64_16x16x4f64: D (16 x 16) = A (16 x 4) B (4 x 16) + C.  Lane l of the
// wave supplies A[l % 16][l / 16] and B[l / 16][l % 16] and holds
// D[l / 16 + 4 r][l % 16] in component r.  All 64 lanes take part.
typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dbl4 mfma_f64_16x16x4(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}


// One product on the matrix cores for all elements of a wave:
// OUT[c][i] = sum_j M[i][j] IN[c][j], i < ROWS, j < KD, for the columns
// c = t NFLD + f (element t < TPW of the wave, field f).  M is column major
// (M[i + j LDM]), in global memory or, with M_LDS, in LDS; IN and OUT are
// per-element LDS arrays, element t's field f at smem + t PER + in_off + f LDIN
// and smem + t PER + out_off + f LDOUT.  The product is padded to 16-row,
// 16-column and 4-deep tiles with zeros (elements at or beyond nvalid give zero
// columns, which are not stored).  Each column's sum runs over j in ascending
// order, one fused multiply-add per term, as the vector-ALU row products do (the
// results are the same bits).  Every lane of the wave must call this, and IN
// must not overlap OUT.
template <int NFLD, int TPW, int PER, int ROWS, int KD, int LDM, int LDIN, int LDOUT,
          bool M_LDS = false>
__device__ __forceinline__ void wave_product_mfma(const double *M, int in_off, int out_off,
                                                  double *smem, int lane, int nvalid) {
  constexpr int NCOL = TPW * NFLD;
  constexpr int MT = (ROWS + 15) / 16, NT = (NCOL + 15) / 16, KS = (KD + 3) / 4;
  const int lr = lane & 15, lk = lane >> 4;
  dbl4 acc[MT][NT];
  for (int mt = 0; mt < MT; ++mt)
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = dbl4{0.0, 0.0, 0.0, 0.0};
  int boff[NT], ooff[NT];
  bool bok[NT];
  for (int nt = 0; nt < NT; ++nt) {
    const int c = 16 * nt + lr;
    const int t = c / NFLD;
    const int f = c - t * NFLD;
    bok[nt] = c < NCOL && t < nvalid;
    boff[nt] = bok[nt] ? t * PER + in_off + f * LDIN : 0;
    ooff[nt] = bok[nt] ? t * PER + out_off + f * LDOUT : 0;
  }
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int j = 4 * ks + lk;
    double a[MT], b[NT];
    for (int mt = 0; mt < MT; ++mt) {
      const int row = 16 * mt + lr;
      if (j < KD && row < ROWS)
        a[mt] = M_LDS ? M[row + j * LDM] : ld_off(M, (unsigned)(row + j * LDM));
      else
        a[mt] = 0.0;
    }
    for (int nt = 0; nt < NT; ++nt) b[nt] = (j < KD && bok[nt]) ? smem[boff[nt] + j] : 0.0;
    for (int mt = 0; mt < MT; ++mt)
      for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma_f64_16x16x4(a[mt], b[nt], acc[mt][nt]);
  }
  for (int nt = 0; nt < NT; ++nt) {
    if (!bok[nt]) continue;
    for (int mt = 0; mt < MT; ++mt)
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * mt + lk + 4 * r;
        if (i < ROWS) smem[ooff[nt] + i] = acc[mt][nt][r];
      }
  }
}

// ------------------------------------------------------------------ update kernel

// Launch shape of update_kernel: lane l of an element's LPE lanes holds rows
// l + r * LPE (r < R) of the matrix-vector products, so each value broadcast
// from LDS feeds R rows.  At 3D N = 3 two rows per lane put 5 elements of 12
// lanes in a wave (60 lanes busy) instead of 2 of 23 (46); at 3D N = 4 two
// elements of 22 lanes instead of one of 44.  Each row sums over j in ascending
// order whatever R is, so R changes no result.  VfPq is staged in LDS where the
// workgroup's LDS stays within 64 KB; otherwise the face products read it
// through L1 (consecutive lanes, consecutive face points: contiguous).
template <class S> struct UpdShape {
  static constexpr int R = (S::DIM == 3 && (S::NQ == 23 || S::NQ == 44)) ? 2 : 1;
  static constexpr int LPE = (S::NQ + R - 1) / R;
  static constexpr int TPW = WAVE / LPE;
  static constexpr int EPB = 4 * TPW;   // four waves per workgroup
  static constexpr bool PF_LDS =
      8 * (S::NQ * S::NQ + S::NFN * S::NQ + 2 * EPB * S::NFLD * S::NQ) <= 65536;
};

template <class S>
__global__ void __launch_bounds__(WAVE * 4)
update_kernel(int K, const double *__restrict__ Jq, const double *__restrict__ VqPq,
              const double *__restrict__ VfPq, double fa, double fb, double fdt,
              double *__restrict__ rhsQ, double *__restrict__ resQ, double *__restrict__ Q,
              double *__restrict__ Qf, double gamma, double lgm1) {
  constexpr int DIM = S::DIM, NQ = S::NQ, NFN = S::NFN, NFLD = S::NFLD;
  constexpr int R = UpdShape<S>::R, LPE = UpdShape<S>::LPE, TPW = UpdShape<S>::TPW;
  constexpr int EPB = UpdShape<S>::EPB;
  constexpr bool PF_LDS = UpdShape<S>::PF_LDS;
  __shared__ double sP[NQ * NQ];
  __shared__ double sPf[PF_LDS ? NFN * NQ : 1];
  __shared__ double buf[2][EPB][NFLD][NQ];
  double (*bufA)[NFLD][NQ] = buf[0];
  double (*bufB)[NFLD][NQ] = buf[1];
  for (int idx = threadIdx.x; idx < NQ * NQ; idx += blockDim.x) sP[idx] = VqPq[idx];
  if constexpr (PF_LDS)
    for (int idx = threadIdx.x; idx < NFN * NQ; idx += blockDim.x) sPf[idx] = VfPq[idx];

  const int wave = threadIdx.x / WAVE;
  const int wl = threadIdx.x % WAVE;
  const int tw = wl / LPE;
  const int l = wl - tw * LPE;
  const bool active = tw < TPW;
  const int e = wave * TPW + (active ? tw : 0);
  const int k = blockIdx.x * EPB + e;
  const bool valid = active && k < K;
  const size_t base = (size_t)k * NQ * NFLD;
  // rows of this lane; a row beyond NQ (last lane of an element) is not computed
  int ir[R];
  bool rv[R];
  for (int x = 0; x < R; ++x) {
    ir[x] = l + x * LPE;
    rv[x] = valid && ir[x] < NQ;
    if (!rv[x]) ir[x] = l;
  }
  // out = c VqPq buf[src] for this lane's rows: the wave's elements' product on
  // the matrix cores (left in buf[1 - src]), read back by rows.  All threads
  // call it.
  const int kw = blockIdx.x * EPB + wave * TPW;           // first element of this wave
  const int nvalid = K - kw < TPW ? (K - kw < 0 ? 0 : K - kw) : TPW;
  auto product = [&](int src, double c, double (&out)[R][NFLD]) {
    constexpr int PE = NFLD * NQ, OFF = EPB * NFLD * NQ;
    wave_product_mfma<NFLD, TPW, PE, NQ, NQ, NQ, NQ, NQ, true>(
        sP, src * OFF, (1 - src) * OFF, &buf[0][wave * TPW][0][0], wl, nvalid);
    __syncthreads();
    if (valid)
      for (int x = 0; x < R; ++x)
        for (int f = 0; f < NFLD; ++f) out[x][f] = c * buf[1 - src][e][f][ir[x]];
  };

  double J[R], invJ[R];
  for (int x = 0; x < R; ++x) {
    J[x] = invJ[x] = 1.0;
    if (rv[x]) {
      J[x] = Jq[(size_t)k * NQ + ir[x]];
      invJ[x] = 1.0 / J[x];
      for (int f = 0; f < NFLD; ++f) bufA[e][f][ir[x]] = rhsQ[base + ir[x] + f * NQ];
    }
  }
  __syncthreads();
  // r = -2 VqPq rhs ; bufB = r / J
  {
    double r[R][NFLD];
    product(0, -2.0, r);
    if (valid)
      for (int x = 0; x < R; ++x)
        if (rv[x])
          for (int f = 0; f < NFLD; ++f)
            bufB[e][f][ir[x]] = (DIM == 3) ? r[x][f] * invJ[x] : r[x][f] / J[x];
  }
  __syncthreads();
  // r = VqPq bufB ; RK update ; bufA = v(Q) J
  {
    double r[R][NFLD];
    product(1, 1.0, r);
    if (valid)
      for (int x = 0; x < R; ++x) {
        if (!rv[x]) continue;
        double U[NFLD], V[NFLD];
        for (int f = 0; f < NFLD; ++f) {
          const size_t id = base + ir[x] + f * NQ;
          const double res = fa * resQ[id] + fdt * r[x][f];
          resQ[id] = res;
          U[f] = Q[id] + fb * res;
          Q[id] = U[f];
        }
        entropy_vars<DIM>(U, gamma, V);
        for (int f = 0; f < NFLD; ++f) bufA[e][f][ir[x]] = V[f] * J[x];
      }
  }
  __syncthreads();
  // bufB = VqPq bufA / J
  {
    double r[R][NFLD];
    product(0, 1.0, r);
    if (valid)
      for (int x = 0; x < R; ++x)
        if (rv[x])
          for (int f = 0; f < NFLD; ++f)
            bufB[e][f][ir[x]] = (DIM == 3) ? r[x][f] * invJ[x] : r[x][f] / J[x];
  }
  __syncthreads();
  // bufA = VqPq bufB   (entropy variables v~: the product leaves them in bufA)
  {
    constexpr int PE = NFLD * NQ, OFF = EPB * NFLD * NQ;
    wave_product_mfma<NFLD, TPW, PE, NQ, NQ, NQ, NQ, NQ, true>(
        sP, OFF, 0, &buf[0][wave * TPW][0][0], wl, nvalid);
  }
  __syncthreads();
  if (valid) {
    // face points: Qf = u(VfPq v~)
    const size_t basef = (size_t)k * NFN * NFLD;
    for (int fn = l; fn < NFN; fn += LPE) {
      double V[NFLD], U[NFLD];
      for (int f = 0; f < NFLD; ++f) V[f] = 0.0;
#pragma unroll 2
      for (int j = 0; j < NQ; ++j) {
        double P;
        if constexpr (PF_LDS) P = sPf[fn + j * NFN];
        else P = ld_off(VfPq, (unsigned)(fn + j * NFN));
        for (int f = 0; f < NFLD; ++f) V[f] += P * bufA[e][f][j];
      }
      conservative_vars<DIM>(V, gamma, lgm1, U);
      for (int f = 0; f < NFLD; ++f) Qf[basef + fn + f * NFN] = U[f];
    }
    // volume points: rhsQ = u(VqPq v~) = u(v~) (VqPq is a projection and v~
    // already lies in its range, round-off apart)
    for (int x = 0; x < R; ++x) {
      if (!rv[x]) continue;
      double V[NFLD], U[NFLD];
      for (int f = 0; f < NFLD; ++f) V[f] = bufA[e][f][ir[x]];
      conservative_vars<DIM>(V, gamma, lgm1, U);
      for (int f = 0; f < NFLD; ++f) rhsQ[base + ir[x] + f * NQ] = U[f];
    }
  }
}

// ------------------------------------------------------------------ fused stage kernel
//
// One LSRK stage for one element per team in a single launch: volume term,
// surface term (neighbour traces through mapPq) and its lift, weight-adjusted
// inverse mass, RK update and entropy projection.  The RHS never leaves the
// chip.  Neighbours read face traces of the previous stage while this element
// writes its new ones, so the face traces are double-buffered: Qf_in is read,
// Qf_out written (the caller swaps them between stages).

// Arguments of one fused stage over one mesh; the ensemble kernel keeps one per
// problem in device memory.
struct StageArgs {
  int K;
  const double *vgeo, *vfgeo, *fgeo;
  const uint32_t *mapPq;
  const double *nrJ, *nsJ, *ntJ, *ws, *VqLq, *VqPq, *VfPq, *Jq;
  double fa, fb, fdt;
  const double *Qf_in;
  double *Qf_out, *rhsQ, *resQ, *Q;            // outputs
  const double *rhsQ_in, *resQ_in, *Q_in;      // inputs (equal to the outputs: in place)
  double gamma, gm1, inv2gm1, ltau, lgm1;
};

// Room for the fused stage's [NFLD][NFN] face-trace product behind the
// update's two [NFLD][NQ] buffers.
template <class S> constexpr bool traces_mfma() {
  return VolSmem<S>::per_elem >= 2 * S::NFLD * S::NQ + S::NFLD * S::NFN;
}

// The fused stage's products: elements at the stage's LDS stride.
template <class S, int ROWS, int KD, int LDM, int LDIN, int LDOUT>
__device__ __forceinline__ void stage_product_mfma(const double *M, int in_off, int out_off,
                                                   double *smem, int lane, int nvalid) {
  wave_product_mfma<S::NFLD, S::TPW, VolSmem<S>::per_elem, ROWS, KD, LDM, LDIN, LDOUT>(
      M, in_off, out_off, smem, lane, nvalid);
}

// Stage arguments are read through the constant address space, where they
// live (kernel arguments, or descriptor tables written by the host): loads are
// scalar, can be repeated at each use instead of held in registers, and the
// pointers loaded from there are known to point to global memory.
typedef __attribute__((address_space(4))) const StageArgs ConstArgs;

// One fused stage for the S::TPW elements kfirst .. kfirst+TPW-1 (one team of
// S::LPE lanes each, lane l holding rows l + r * LPE) of the mesh described by
// A.  All lanes of the workgroup call it.  fa, fb and the face-trace buffers
// come separately from A so that an ensemble can keep A in device memory and
// vary them per launch.  lane is threadIdx.x (passed in so that a caller
// looping over tiles can hand an opaque copy).
template <class S>
__device__ __forceinline__ void stage_body(const ConstArgs *A, double fa, double fb,
                                           const double *__restrict__ Qf_in,
                                           double *__restrict__ Qf_out, int kfirst, int lane,
                                           double (*s_nref)[S::DIM], double *smem) {
  constexpr int DIM = S::DIM, NQ = S::NQ, NFN = S::NFN, NFLD = S::NFLD, NG = S::NG;
  constexpr int NFG = DIM + 1, R = S::R, LPE = S::LPE, TPW = S::TPW;
  constexpr int PER = VolSmem<S>::per_elem;
  const double gamma = A->gamma, gm1 = A->gm1, inv2gm1 = A->inv2gm1, ltau = A->ltau, lgm1 = A->lgm1;

  const int team = lane / LPE;
  const int l = lane - team * LPE;
  const bool active = team < TPW;
  const int k = kfirst + team;
  const bool valid = active && k < A->K;
  const size_t kk = valid ? (size_t)k : 0;
  const size_t base = kk * NQ * NFLD;
  const double fdt = A->fdt;
  bool rv[R];                     // row r = l + r * LPE of this lane exists
  for (int r = 0; r < R; ++r) rv[r] = valid && (r == 0 || l + r * LPE < NQ);

  load_nref<S>(lane, A->nrJ, A->nsJ, A->ntJ, s_nref);
  double *accV = smem + (active ? team : 0) * PER;
  double acc[R][NFLD];
  const WaveCtx wc{lane, kfirst, A->K, Qf_in, A->vfgeo, smem};
  vol_body<S>(valid, l, A->rhsQ_in + base, A->vgeo + kk * NQ * NG, s_nref, A->ws, accV, gm1,
              inv2gm1, acc, wc);

  // The update's per-row inputs.  With R > 1 and in 2D they are loaded now, so
  // that their HBM latency passes under the surface and lift phases (2D: the
  // ensemble 3 % and the single-problem stage 4 % faster).  The 3D one-row
  // shapes load them where they are used: N = 2 has no registers to spare, and
  // at N = 4 loading them early measured the same.
  constexpr bool EARLY = R > 1 || DIM == 2;
  int ir[R];
  for (int r = 0; r < R; ++r) ir[r] = rv[r] ? l + r * LPE : 0;
  double Qin[EARLY ? R : 1][NFLD], resin[EARLY ? R : 1][NFLD], J[R];
  for (int r = 0; r < R; ++r) {
    J[r] = 1.0;
    if (EARLY && rv[r]) {
      J[r] = A->Jq[kk * NQ + ir[r]];
      for (int f = 0; f < NFLD; ++f) {
        const size_t id = base + ir[r] + f * NQ;
        Qin[r][f] = A->Q_in[id];
        resin[r][f] = A->resQ_in[id];
      }
    }
  }

  // ---------------- surface flux + LF penalty, in place over the face rows.
  // The face points of all TPW elements are spread over the whole wave.
  for (int p = lane; p < TPW * NFN; p += WAVE) {
    const int t = p / NFN;
    const int fn = p - t * NFN;
    const int kt = kfirst + t;
    if (kt >= A->K) continue;
    const size_t kq = (size_t)kt;
    double *accFt = smem + t * PER + VolSmem<S>::accF_off;
    const double *g = A->fgeo + kq * NFN * NFG;
    const size_t idM = (size_t)fn + kq * NFN * NFLD;
    const size_t idP = A->mapPq[kq * NFN + fn];
    double UM[NFLD], UP[NFLD];
    for (int f = 0; f < NFLD; ++f) {
      UM[f] = Qf_in[idM + (size_t)f * NFN];
      UP[f] = Qf_in[idP + (size_t)f * NFN];
    }
    double nJ[DIM];
    for (int d = 0; d < DIM; ++d) nJ[d] = g[fn + d * NFN];
    const double sJ = g[fn + DIM * NFN];
    // 3D: the logarithms are taken only if a lane of the wave needs the log branch
    const Node<DIM> a = make_node<DIM, DIM == 2>(UM, gm1);
    const Node<DIM> b = make_node<DIM, DIM == 2>(UP, gm1);
    const double pM = gm1 * (UM[NFLD - 1] - 0.5 * UM[0] * (2.0 * a.hq));
    const double pP = gm1 * (UP[NFLD - 1] - 0.5 * UP[0] * (2.0 * b.hq));
    const double lamM = sqrt(2.0 * a.hq) + sqrt(gamma * pM / UM[0]);
    const double lamP = sqrt(2.0 * b.hq) + sqrt(gamma * pP / UP[0]);
    const double Lfc = ltau * fmax(lamM, lamP) * sJ;
    const PairFlux<DIM> P = pair_flux<DIM, DIM == 3, (S::R > 1)>(a, b, inv2gm1);
    double Fn[NFLD];
    flux_dot<DIM>(P, nJ, Fn);
    for (int f = 0; f < NFLD; ++f)
      accFt[f * NFN + fn] = 0.5 * Fn[f] + accFt[f * NFN + fn] - Lfc * (UP[f] - UM[f]);
  }
  __syncthreads();
  // ---------------- RHS rows = volume part + lift of the face terms
  const int nvalid = A->K - kfirst < TPW ? A->K - kfirst : TPW;   // elements of this wave
  static_assert(VolSmem<S>::accF_off >= 2 * NFLD * NQ, "the lift's result goes behind bufA");
  double rhs[R][NFLD];
  // the lift of all the wave's elements as one matrix-core product, into the
  // face-record region (free by now)
  stage_product_mfma<S, NQ, NFN, NQ, NFN, NQ>(A->VqLq, VolSmem<S>::accF_off, NFLD * NQ, smem, lane,
                                             nvalid);
  __syncthreads();
  if (valid)
    for (int r = 0; r < R; ++r)
      for (int f = 0; f < NFLD; ++f)
        rhs[r][f] = (acc[r][f] + accV[f * NQ + ir[r]]) + accV[NFLD * NQ + f * NQ + ir[r]];
  __syncthreads();

  // ---------------- update (same sequence as update_kernel)
  double *bufA = accV;                // [NFLD][NQ]
  double *bufB = accV + NFLD * NQ;    // [NFLD][NQ]
  double invJ[R];
  for (int r = 0; r < R; ++r) {
    invJ[r] = 1.0;
    if (rv[r]) {
      if (!EARLY) J[r] = A->Jq[kk * NQ + ir[r]];
      invJ[r] = 1.0 / J[r];
      for (int f = 0; f < NFLD; ++f) bufA[f * NQ + ir[r]] = rhs[r][f];
    }
  }
  __syncthreads();
  // The four VqPq products: on the matrix cores for all the wave's elements at
  // once, each written to LDS and read back by rows.  res = scale VqPq (buffer
  // at in_off) for this lane's rows; the product also stays in the buffer at
  // out_off.  All lanes call it.
  auto product = [&](int in_off, int out_off, double scale, double (&res)[R][NFLD]) {
    stage_product_mfma<S, NQ, NQ, NQ, NQ, NQ>(A->VqPq, in_off, out_off, smem, lane, nvalid);
    __syncthreads();
    if (valid)
      for (int r = 0; r < R; ++r)
        for (int f = 0; f < NFLD; ++f) res[r][f] = scale * accV[out_off + f * NQ + ir[r]];
  };
  constexpr int OA = 0, OB = NFLD * NQ;   // bufA, bufB
  {
    double res[R][NFLD];
    product(OA, OB, -2.0, res);
    if (valid)
      for (int r = 0; r < R; ++r) {
        if (!rv[r]) continue;
        for (int f = 0; f < NFLD; ++f)
          bufB[f * NQ + ir[r]] = (DIM == 3) ? res[r][f] * invJ[r] : res[r][f] / J[r];
      }
  }
  __syncthreads();
  {
    double res[R][NFLD];
    product(OB, OA, 1.0, res);
    if (valid)
      for (int r = 0; r < R; ++r) {
        if (!rv[r]) continue;
        double U[NFLD], V[NFLD];
        for (int f = 0; f < NFLD; ++f) {
          const size_t id = base + ir[r] + f * NQ;
          const double rq = fa * (EARLY ? resin[r][f] : A->resQ_in[id]) + fdt * res[r][f];
          A->resQ[id] = rq;
          U[f] = (EARLY ? Qin[r][f] : A->Q_in[id]) + fb * rq;
          A->Q[id] = U[f];
        }
        entropy_vars<DIM>(U, gamma, V);
        for (int f = 0; f < NFLD; ++f) bufA[f * NQ + ir[r]] = V[f] * J[r];
      }
  }
  __syncthreads();
  {
    double res[R][NFLD];
    product(OA, OB, 1.0, res);
    if (valid)
      for (int r = 0; r < R; ++r) {
        if (!rv[r]) continue;
        for (int f = 0; f < NFLD; ++f)
          bufB[f * NQ + ir[r]] = (DIM == 3) ? res[r][f] * invJ[r] : res[r][f] / J[r];
      }
  }
  __syncthreads();
  // entropy variables v~ into bufA
  stage_product_mfma<S, NQ, NQ, NQ, NQ, NQ>(A->VqPq, OB, OA, smem, lane, nvalid);
  __syncthreads();
  // face points, spread over the whole wave: Qf = u(VfPq v~), with VfPq v~ of
  // all the wave's elements as one matrix-core product
  static_assert(traces_mfma<S>(), "the traces' entropy variables go behind bufB");
  constexpr int WOFF = 2 * NFLD * NQ;
  stage_product_mfma<S, NFN, NQ, NFN, NQ, NFN>(A->VfPq, 0, WOFF, smem, lane, nvalid);
  __syncthreads();
  for (int p = lane; p < TPW * NFN; p += WAVE) {
    const int t = p / NFN;
    const int fn = p - t * NFN;
    const int kt = kfirst + t;
    if (kt >= A->K) continue;
    const double *bufAt = smem + t * PER;
    double V[NFLD], U[NFLD];
    for (int f = 0; f < NFLD; ++f) V[f] = bufAt[WOFF + f * NFN + fn];
    conservative_vars<DIM>(V, gamma, lgm1, U);
    double *qf = Qf_out + (size_t)kt * NFN * NFLD;
    for (int f = 0; f < NFLD; ++f) qf[fn + f * NFN] = U[f];
  }
  // volume points: rhsQ = u(VqPq v~) = u(v~), since VqPq is a projection and
  // v~ = VqPq (...) already lies in its range (round-off apart)
  for (int r = 0; r < R; ++r) {
    if (!rv[r]) continue;
    double V[NFLD], U[NFLD];
    for (int f = 0; f < NFLD; ++f) V[f] = bufA[f * NQ + ir[r]];
    conservative_vars<DIM>(V, gamma, lgm1, U);
    for (int f = 0; f < NFLD; ++f) A->rhsQ[base + ir[r] + f * NQ] = U[f];
  }
}

// ------------------------------------------------------------------ diagnostics
//
// Per element: aux[k][0] = sum wq J |m|^2/(2 rho)      (euler_3d_compute_aux, okl/Euler3D.okl:3888-3934)
//              aux[k][1] = sum wq J (-rho s)           (entropy, euler.c:4135-4156)
//              aux[k][2] = sum wq J |Q - Qex|^2        (squared L2 error, euler.c:4227-4274)
//              aux[k][3] = sum wq J rho                (mass)
// then totals[c] = sum_k aux[k][c] by one workgroup in a fixed order.
constexpr int NDIAG = 4;

template <class S, int TPW>
__global__ void __launch_bounds__(WAVE)
diag_kernel(int K, const double *__restrict__ wq, const double *__restrict__ Jq,
            const double *__restrict__ Q, const double *__restrict__ Qex, double *__restrict__ aux,
            double gamma) {
  constexpr int DIM = S::DIM, NQ = S::NQ, NFLD = S::NFLD;
  __shared__ double part[TPW][NDIAG][NQ];
  const int lane = threadIdx.x;
  const int team = lane / NQ;
  const int i = lane - team * NQ;
  const bool active = team < TPW;
  const int k = blockIdx.x * TPW + team;
  const bool valid = active && k < K;
  if (valid) {
    const size_t id = (size_t)k * NQ * NFLD + i;
    double U[NFLD];
    for (int f = 0; f < NFLD; ++f) U[f] = Q[id + f * NQ];
    const double wJ = wq[i] * Jq[(size_t)k * NQ + i];
    double mm = U[1] * U[1];
    for (int d = 1; d < DIM; ++d) mm = mm + U[1 + d] * U[1 + d];
    const double rhoe = U[NFLD - 1] - 0.5 * mm / U[0];
    const double s = log((gamma - 1.0) * rhoe) - gamma * log(U[0]);
    double err = 0.0;
    if (Qex) {
      for (int f = 0; f < NFLD; ++f) {
        const double e = U[f] - Qex[id + f * NQ];
        err = err + e * e;
      }
    }
    part[team][0][i] = 0.5 * mm / U[0] * wJ;
    part[team][1][i] = wJ * (-U[0] * s);
    part[team][2][i] = err * wJ;
    part[team][3][i] = wJ * U[0];
  }
  __syncthreads();
  if (valid && i < NDIAG) {
    double t = 0.0;
    for (int j = 0; j < NQ; ++j) t += part[team][i][j];
    aux[(size_t)k * NDIAG + i] = t;
  }
}

__global__ void __launch_bounds__(1024)
diag_reduce_kernel(int K, const double *__restrict__ aux, double *__restrict__ totals) {
  __shared__ double red[NDIAG][1024];
  const int t = threadIdx.x;
  double acc[NDIAG] = {0.0, 0.0, 0.0, 0.0};
  for (int k = t; k < K; k += 1024)
    for (int c = 0; c < NDIAG; ++c) acc[c] += aux[(size_t)k * NDIAG + c];
  for (int c = 0; c < NDIAG; ++c) red[c][t] = acc[c];
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w)
      for (int c = 0; c < NDIAG; ++c) red[c][t] += red[c][t + w];
    __syncthreads();
  }
  if (t < NDIAG) totals[t] = red[t][0];
}


template <class S>
__global__ void __launch_bounds__(WAVE)
__attribute__((amdgpu_waves_per_eu(vol_waves_min<S>(), vol_waves<S>()))) stage_kernel(StageArgs A) {
  extern __shared__ double smem[];
  __shared__ double s_nref[S::NFACES][S::DIM];
  stage_body<S>((const ConstArgs *)&A, A.fa, A.fb, A.Qf_in, A.Qf_out, blockIdx.x * S::TPW,
                threadIdx.x, s_nref, smem);
}

#include "wadges_adjoint.hpp"

// ------------------------------------------------------------------ dispatch

thread_local int g_last_hip = 0;

inline int hip_check(hipError_t e) {
  if (e != hipSuccess) {
    g_last_hip = (int)e;
    return WADGES_EHIP;
  }
  return WADGES_OK;
}

// Per-element partials of wadges_diagnostics_noalloc: a fixed array of the
// library, one copy per GPU, so that call allocates nothing.  Host code never
// passes `scratch` itself to a kernel (on the host the name is only a stand-in
// for the symbol): it passes the device address that hipGetSymbolAddress
// returns for the current device, looked up once per device.
static __device__ __align__(256) double scratch[WADGES_NOALLOC_MAX_K * NDIAG];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be the current device.
inline int scratch_address(int device, double **addr) {
  if (device < 0 || device >= kMaxDevices) return WADGES_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void *p = nullptr;
    const int rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (rc) return rc;
    g_scratch_addr[device] = (double *)p;
  }
  *addr = g_scratch_addr[device];
  return WADGES_OK;
}

struct Ops {
  int (*vol)(const wadges_plan *, int, const double *, const double *, const double *,
             const double *, const double *, const double *, const double *, const double *,
             const double *, const double *, const double *, double *, double *, hipStream_t);
  int (*surf)(const wadges_plan *, int, const double *, const uint32_t *, const double *,
              const double *, const double *, double *, hipStream_t);
  int (*update)(const wadges_plan *, int, const double *, const double *, const double *,
                double, double, double, double *, double *, double *, double *, hipStream_t);
  int (*stage)(const wadges_plan *, int, const double *, const double *, const double *,
               const uint32_t *, const double *, const double *, const double *, const double *,
               const double *, const double *, const double *, const double *, const double *,
               const double *, double, double, double, const double *, double *,
               const double *, double *, const double *, double *, const double *, double *,
               hipStream_t);
  int (*stage_vjp)(const wadges_plan *, int, const double *, const double *, const double *,
                   const uint32_t *, const double *, const double *, const double *,
                   const double *, const double *, const double *, const double *,
                   const double *, const double *, const double *, double, double, double,
                   const double *, const double *, const double *, double *, double *,
                   const double *, const double *, double *, double *, double *, double *,
                   hipStream_t);
  int (*diag)(const wadges_plan *, int, const double *, const double *, const double *,
              const double *, double *, double *, hipStream_t);
  int (*rhs_vjp)(const wadges_plan *, int, const double *, const double *, const double *,
                 const uint32_t *, const double *, const double *, const double *, const double *,
                 const double *, const double *, const double *, const double *, const double *,
                 const double *, const double *, double *, double *, double *, hipStream_t);
  int (*update_vjp)(const wadges_plan *, int, const double *, const double *, const double *,
                    double, double, double, const double *, double *, double *, const double *,
                    const double *, double *, hipStream_t);
  int (*proj_vjp)(const wadges_plan *, int, const double *, const double *, const double *,
                  const double *, const double *, double *, hipStream_t);
  int (*prep)(const double *, const double *, const double *, const double *, const double *,
              double *, hipStream_t);
  int ws_doubles;
  int stage_smem;  // dynamic LDS bytes of one fused-stage workgroup
  int stage_tpw;   // elements per fused-stage workgroup
};

}  // namespace wadges

struct wadges_plan {
  int dim, N, Nq, Nfq, Nfaces, Nf, nfields;
  double gamma, tau;
  double *ws;  // device workspace: two sets of operator tables
  int device;
  // Device address of the library's fixed diagnostics array on this plan's
  // device (wadges_diagnostics_noalloc).
  double *scratch_addr;
  wadges::Ops ops;
  // Operator pointers the tables of the calls outside capture (ws[0, WS)) were
  // prepared from (null: not prepared).
  mutable const double *tab_src[5];
  // The capture (its id) that last recorded a preparation of the captured
  // tables (ws[WS, 2 WS)), and from which operator pointers.
  mutable unsigned long long cap_id;
  mutable const double *cap_src[5];
};

namespace wadges {

// Prepares the operator tables of the plan unless they were prepared from the
// same operator pointers before, and returns in *tab the tables the kernels of
// this call read.  The plan holds two sets: one for calls outside stream
// capture and one for captured calls.  Outside capture a fresh preparation is
// waited for, so the tables are then valid for calls on every stream.  Under
// capture the preparation of the captured set is recorded into the graph at
// the capture's first call with these operators: every graph rewrites the
// tables it reads whenever it runs, and no graph touches the tables of the
// calls outside capture.
template <class S>
int ensure_tables(const wadges_plan *p, const double *Drq, const double *Dsq, const double *Dtq,
                  const double *VqLq, const double *VfPq, hipStream_t st, const double **tab) {
  const double *src[5] = {Drq, Dsq, Dtq, VqLq, VfPq};
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  unsigned long long id = 0;
  int rc = hip_check(hipStreamGetCaptureInfo(st, &cs, &id));
  if (rc) return rc;
  const bool capturing = cs != hipStreamCaptureStatusNone;
  double *ws = capturing ? p->ws + S::WS : p->ws;
  *tab = ws;
  if (capturing ? (id == p->cap_id && std::equal(src, src + 5, p->cap_src))
                : std::equal(src, src + 5, p->tab_src))
    return WADGES_OK;
  prep_vol_tables<S><<<(S::WS + 255) / 256, 256, 0, st>>>(Drq, Dsq, Dtq, VqLq, VfPq, ws);
  if ((rc = hip_check(hipGetLastError()))) return rc;
  if (capturing) {
    p->cap_id = id;
    std::copy(src, src + 5, p->cap_src);
    return WADGES_OK;
  }
  if ((rc = hip_check(hipStreamSynchronize(st)))) return rc;
  std::copy(src, src + 5, p->tab_src);
  return WADGES_OK;
}

template <class S> int vol_impl(const wadges_plan *p, int K, const double *vgeo,
                                const double *vfgeo, const double *nrJ, const double *nsJ,
                                const double *ntJ, const double *Drq, const double *Dsq,
                                const double *Dtq, const double *VqLq, const double *VfPq,
                                const double *Qf, double *rhsQ, double *rhsQf, hipStream_t st) {
  if (K == 0) return WADGES_OK;
  const double *tab = nullptr;
  {
    int rc = ensure_tables<S>(p, Drq, Dsq, Dtq, VqLq, VfPq, st, &tab);
    if (rc) return rc;
  }
  constexpr int TPW = S::TPW;
  static_assert(TPW >= 1, "Nq must fit in one wave");
  const size_t smem = (size_t)TPW * VolSmem<S>::per_elem * sizeof(double);
  const int blocks = (K + TPW - 1) / TPW;
  vol_kernel<S><<<blocks, WAVE, smem, st>>>(K, vgeo, vfgeo, nrJ, nsJ, ntJ, tab, Qf, rhsQ,
                                                 rhsQf, p->gamma - 1.0,
                                                 1.0 / (2.0 * (p->gamma - 1.0)));
  return hip_check(hipGetLastError());
}

template <class S> int surf_impl(const wadges_plan *p, int K, const double *fgeo,
                                 const uint32_t *mapPq, const double *VqLq, const double *Qf,
                                 const double *rhsQf, double *rhsQ, hipStream_t st) {
  if (K == 0) return WADGES_OK;
  constexpr int EPB = surf_epb(S::NFN);
  const int blocks = (K + EPB - 1) / EPB;
  surf_kernel<S, EPB><<<blocks, WAVE * 4, 0, st>>>(K, fgeo, mapPq, VqLq, Qf, rhsQf, rhsQ,
                                                   p->gamma, p->gamma - 1.0,
                                                   1.0 / (2.0 * (p->gamma - 1.0)),
                                                   0.25 * p->tau);
  return hip_check(hipGetLastError());
}

template <class S> int update_impl(const wadges_plan *p, int K, const double *Jq,
                                   const double *VqPq, const double *VfPq, double fa, double fb,
                                   double fdt, double *rhsQ, double *resQ, double *Q, double *Qf,
                                   hipStream_t st) {
  if (K == 0) return WADGES_OK;
  constexpr int EPB = UpdShape<S>::EPB;
  const int blocks = (K + EPB - 1) / EPB;
  update_kernel<S><<<blocks, WAVE * 4, 0, st>>>(K, Jq, VqPq, VfPq, fa, fb, fdt, rhsQ, resQ, Q, Qf,
                                                p->gamma, std::log(p->gamma - 1.0));
  return hip_check(hipGetLastError());
}

template <class S> int stage_impl(const wadges_plan *p, int K, const double *vgeo,
                                  const double *vfgeo, const double *fgeo, const uint32_t *mapPq,
                                  const double *nrJ, const double *nsJ, const double *ntJ,
                                  const double *Drq, const double *Dsq, const double *Dtq,
                                  const double *VqLq, const double *VqPq, const double *VfPq,
                                  const double *Jq, double fa, double fb, double fdt,
                                  const double *Qf_in, double *Qf_out, const double *rhsQ_in,
                                  double *rhsQ, const double *resQ_in, double *resQ,
                                  const double *Q_in, double *Q, hipStream_t st) {
  if (K == 0) return WADGES_OK;
  const double *tab = nullptr;
  int rc = ensure_tables<S>(p, Drq, Dsq, Dtq, VqLq, VfPq, st, &tab);
  if (rc) return rc;
  constexpr int TPW = S::TPW;
  const size_t smem = (size_t)TPW * VolSmem<S>::per_elem * sizeof(double);
  const int blocks = (K + TPW - 1) / TPW;
  const double g = p->gamma;
  StageArgs A{K,      vgeo,   vfgeo,   fgeo,   mapPq,  nrJ,     nsJ,  ntJ,
              tab,    VqLq,   VqPq,    VfPq,   Jq,     fa,      fb,   fdt,
              Qf_in,  Qf_out, rhsQ,    resQ,   Q,      rhsQ_in, resQ_in, Q_in,
              g,      g - 1.0, 1.0 / (2.0 * (g - 1.0)), 0.25 * p->tau, std::log(g - 1.0)};
  stage_kernel<S><<<blocks, WAVE, smem, st>>>(A);
  return hip_check(hipGetLastError());
}

template <class S> int diag_impl(const wadges_plan *p, int K, const double *wq, const double *Jq,
                                 const double *Q, const double *Qex, double *aux, double *totals,
                                 hipStream_t st) {
  constexpr int TPW = WAVE / S::NQ;
  if (K > 0) {
    diag_kernel<S, TPW><<<(K + TPW - 1) / TPW, WAVE, 0, st>>>(K, wq, Jq, Q, Qex, aux, p->gamma);
    int rc = hip_check(hipGetLastError());
    if (rc) return rc;
  }
  diag_reduce_kernel<<<1, 1024, 0, st>>>(K, aux, totals);
  return hip_check(hipGetLastError());
}

template <class S> int rhs_vjp_impl(const wadges_plan *p, int K, const double *vgeo,
                                    const double *vfgeo, const double *fgeo, const uint32_t *mapPq,
                                    const double *nrJ, const double *nsJ, const double *ntJ,
                                    const double *Drq, const double *Dsq, const double *Dtq,
                                    const double *VqLq, const double *VfPq, const double *Uq,
                                    const double *Uf, const double *Rb, double *sbar, double *uqb,
                                    double *ufb, hipStream_t st) {
  if (K == 0) return WADGES_OK;
  const double *tab = nullptr;
  int rc = ensure_tables<S>(p, Drq, Dsq, Dtq, VqLq, VfPq, st, &tab);
  if (rc) return rc;
  {
    constexpr int TS = Max<S::NQ, S::NFN>::value;
    constexpr int TPW = WAVE / TS;
    constexpr int EPB = 4 * TPW;
    sbar_kernel<S, TS, TPW><<<(K + EPB - 1) / EPB, WAVE * 4, 0, st>>>(K, VqLq, Rb, sbar);
    rc = hip_check(hipGetLastError());
    if (rc) return rc;
  }
  constexpr int TPW = WAVE / S::NQ;
  const size_t smem = (size_t)TPW * VjpSmem<S>::per_elem * sizeof(double);
  const double g = p->gamma;
  rhs_vjp_kernel<S, TPW><<<(K + TPW - 1) / TPW, WAVE, smem, st>>>(
      K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, tab, Uq, Uf, Rb, sbar, uqb, ufb, g, g - 1.0,
      1.0 / (2.0 * (g - 1.0)), 0.25 * p->tau);
  return hip_check(hipGetLastError());
}

template <class S> int update_vjp_impl(const wadges_plan *p, int K, const double *Jq,
                                       const double *VqPq, const double *VfPq, double fa, double fb,
                                       double fdt, const double *Q, double *Qb, double *resb,
                                       const double *uqb, const double *ufb, double *Rb,
                                       hipStream_t st) {
  if (K == 0) return WADGES_OK;
  constexpr int TPW = WAVE / S::NQ;
  constexpr int EPB = uvjp_wg_waves<S>() * TPW;
  update_vjp_kernel<S, TPW><<<(K + EPB - 1) / EPB, WAVE * uvjp_wg_waves<S>(), 0, st>>>(
      K, Jq, VqPq, VfPq, fa, fb, fdt, Q, Qb, resb, uqb, ufb, Rb, p->gamma, std::log(p->gamma - 1.0),
      nullptr, nullptr);
  return hip_check(hipGetLastError());
}

// VJP of one fused stage: update VJP with the transposed lift (s_bar) fused in,
// then the RHS VJP at the stage's input projected state.  Two launches.
template <class S> int stage_vjp_impl(const wadges_plan *p, int K, const double *vgeo,
                                      const double *vfgeo, const double *fgeo,
                                      const uint32_t *mapPq, const double *nrJ, const double *nsJ,
                                      const double *ntJ, const double *Drq, const double *Dsq,
                                      const double *Dtq, const double *VqLq, const double *VqPq,
                                      const double *VfPq, const double *Jq, double fa, double fb,
                                      double fdt, const double *rhsQ_in, const double *Qf_in,
                                      const double *Q_out, double *Qb, double *resb,
                                      const double *uqb_out, const double *ufb_out, double *Rb,
                                      double *sbar, double *uqb_in, double *ufb_in,
                                      hipStream_t st) {
  if (K == 0) return WADGES_OK;
  const double *tab = nullptr;
  int rc = ensure_tables<S>(p, Drq, Dsq, Dtq, VqLq, VfPq, st, &tab);
  if (rc) return rc;
  const double g = p->gamma;
  {
    constexpr int TPW = WAVE / S::NQ;
    constexpr int EPB = uvjp_wg_waves<S>() * TPW;
    update_vjp_kernel<S, TPW><<<(K + EPB - 1) / EPB, WAVE * uvjp_wg_waves<S>(), 0, st>>>(
        K, Jq, VqPq, VfPq, fa, fb, fdt, Q_out, Qb, resb, uqb_out, ufb_out, Rb, g,
        std::log(g - 1.0), VqLq, sbar);
    if ((rc = hip_check(hipGetLastError()))) return rc;
  }
  constexpr int TPW = WAVE / S::NQ;
  const size_t smem = (size_t)TPW * VjpSmem<S>::per_elem * sizeof(double);
  rhs_vjp_kernel<S, TPW><<<(K + TPW - 1) / TPW, WAVE, smem, st>>>(
      K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, tab, rhsQ_in, Qf_in, Rb, sbar, uqb_in, ufb_in,
      g, g - 1.0, 1.0 / (2.0 * (g - 1.0)), 0.25 * p->tau);
  return hip_check(hipGetLastError());
}

template <class S> int proj_vjp_impl(const wadges_plan *p, int K, const double *VqPq,
                                     const double *VfPq, const double *Q, const double *uqb,
                                     const double *ufb, double *Qb, hipStream_t st) {
  if (K == 0) return WADGES_OK;
  constexpr int TPW = WAVE / S::NQ;
  constexpr int EPB = 4 * TPW;
  proj_vjp_kernel<S, TPW><<<(K + EPB - 1) / EPB, WAVE * 4, 0, st>>>(
      K, VqPq, VfPq, Q, uqb, ufb, Qb, p->gamma, std::log(p->gamma - 1.0));
  return hip_check(hipGetLastError());
}

template <class S> int prep_impl(const double *Drq, const double *Dsq, const double *Dtq,
                                 const double *VqLq, const double *VfPq, double *ws,
                                 hipStream_t st) {
  prep_vol_tables<S><<<(S::WS + 255) / 256, 256, 0, st>>>(Drq, Dsq, Dtq, VqLq, VfPq, ws);
  return hip_check(hipGetLastError());
}

template <class S> constexpr int stage_smem_bytes() {
  return S::TPW * VolSmem<S>::per_elem * (int)sizeof(double);
}

struct Instance {
  int dim, N, Nq, Nfq;
  Ops ops;
};

template <int DIM, int NQ, int NFQ> constexpr Instance make_instance(int N) {
  using S = Shape<DIM, NQ, NFQ>;
  return Instance{DIM, N, NQ, NFQ,
                  Ops{&vol_impl<S>, &surf_impl<S>, &update_impl<S>, &stage_impl<S>,
                      &stage_vjp_impl<S>, &diag_impl<S>,
                      &rhs_vjp_impl<S>, &update_vjp_impl<S>, &proj_vjp_impl<S>, &prep_impl<S>,
                      S::WS, stage_smem_bytes<S>(), S::TPW}};
}

// (dim, N, Nq, Nfq): quadrature of degree 2N, the point counts of the reference's
// tabulated rules (node_data.h) for M = 2N.
const Instance kInstances[] = {
    make_instance<2, 3, 2>(1),  make_instance<2, 6, 3>(2),  make_instance<2, 12, 4>(3),
    make_instance<2, 16, 5>(4), make_instance<2, 25, 6>(5), make_instance<3, 4, 3>(1),
    make_instance<3, 11, 6>(2), make_instance<3, 23, 12>(3), make_instance<3, 44, 16>(4),
};
constexpr int kNumInstances = sizeof(kInstances) / sizeof(kInstances[0]);

}  // namespace wadges

#include "wadges_batch.hpp"

// ================================================================== C ABI

using namespace wadges;

extern "C" {

int wadges_last_hip_error(void) { return g_last_hip; }

const char *wadges_error_string(int code) {
  switch (code) {
    case WADGES_OK: return "ok";
    case WADGES_EINVAL: return "invalid argument";
    case WADGES_EUNSUPPORTED: return "no kernel instance for (dim, Nq, Nfq)";
    case WADGES_EHIP: return "HIP runtime error";
    case WADGES_ENOMEM: return "device allocation failed";
    case WADGES_ETOOLARGE: return "size beyond the fixed capacity of the call";
    default: return "unknown error";
  }
}

int wadges_instance_count(void) { return kNumInstances; }

int wadges_instance(int idx, int *dim, int *N, int *Nq, int *Nfq) {
  if (idx < 0 || idx >= kNumInstances) return WADGES_EINVAL;
  const Instance &I = kInstances[idx];
  if (dim) *dim = I.dim;
  if (N) *N = I.N;
  if (Nq) *Nq = I.Nq;
  if (Nfq) *Nfq = I.Nfq;
  return WADGES_OK;
}

int wadges_plan_create(wadges_plan **plan, int dim, int N, int Nq, int Nfq, double gamma,
                       double tau) {
  if (!plan) return WADGES_EINVAL;
  *plan = nullptr;
  if (!(gamma > 1.0) || !(tau >= 0.0)) return WADGES_EINVAL;
  const Instance *inst = nullptr;
  for (int n = 0; n < kNumInstances; ++n)
    if (kInstances[n].dim == dim && kInstances[n].Nq == Nq && kInstances[n].Nfq == Nfq)
      inst = &kInstances[n];
  if (!inst) return WADGES_EUNSUPPORTED;
  wadges_plan *p = new (std::nothrow) wadges_plan;
  if (!p) return WADGES_ENOMEM;
  p->dim = dim;
  p->N = N;
  p->Nq = Nq;
  p->Nfq = Nfq;
  p->Nfaces = dim + 1;
  p->Nf = Nfq * (dim + 1);
  p->nfields = dim + 2;
  p->gamma = gamma;
  p->tau = tau;
  p->ops = inst->ops;
  p->ws = nullptr;
  for (auto &t : p->tab_src) t = nullptr;
  p->cap_id = 0;
  for (auto &t : p->cap_src) t = nullptr;
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return WADGES_EHIP;
  }
  // the fixed diagnostics array of this device (looked up once per device; a
  // failure here is reported by wadges_diagnostics_noalloc, which retries)
  p->scratch_addr = nullptr;
  if (scratch_address(p->device, &p->scratch_addr) != WADGES_OK) p->scratch_addr = nullptr;
  hipError_t e = hipMalloc(&p->ws, sizeof(double) * 2 * (size_t)inst->ops.ws_doubles);
  if (e != hipSuccess) {
    g_last_hip = (int)e;
    delete p;
    return WADGES_ENOMEM;
  }
  *plan = p;
  return WADGES_OK;
}

int wadges_plan_destroy(wadges_plan *plan) {
  if (!plan) return WADGES_OK;
  int rc = WADGES_OK;
  if (plan->ws) rc = hip_check(hipFree(plan->ws));
  delete plan;
  return rc;
}

int wadges_plan_invalidate(wadges_plan *plan) {
  if (!plan) return WADGES_EINVAL;
  for (auto &t : plan->tab_src) t = nullptr;
  plan->cap_id = 0;
  for (auto &t : plan->cap_src) t = nullptr;
  return WADGES_OK;
}

int wadges_vol(const wadges_plan *p, int K, const double *vgeo, const double *vfgeo,
               const double *nrJ, const double *nsJ, const double *ntJ, const double *Drq,
               const double *Dsq, const double *Dtq, const double *Drstq, const double *VqLq,
               const double *VfPq, const double *Q, const double *Qf, double *rhsQ,
               double *rhsQf, void *stream) {
  (void)Drstq;
  (void)Q;
  if (!p || K < 0) return WADGES_EINVAL;
  if (K > 0 && (!vgeo || !vfgeo || !nrJ || !nsJ || !Drq || !Dsq || !VqLq || !VfPq || !Qf ||
                !rhsQ || !rhsQf))
    return WADGES_EINVAL;
  if (K > 0 && p->dim == 3 && (!ntJ || !Dtq)) return WADGES_EINVAL;
  return p->ops.vol(p, K, vgeo, vfgeo, nrJ, nsJ, ntJ, Drq, Dsq, Dtq, VqLq, VfPq, Qf, rhsQ, rhsQf,
                    (hipStream_t)stream);
}

int wadges_surf(const wadges_plan *p, int K, const double *fgeo, const uint32_t *mapPq,
                const double *VqLq, const double *Qf, const double *rhsQf, double *rhsQ,
                void *stream) {
  if (!p || K < 0) return WADGES_EINVAL;
  if (K > 0 && (!fgeo || !mapPq || !VqLq || !Qf || !rhsQf || !rhsQ)) return WADGES_EINVAL;
  return p->ops.surf(p, K, fgeo, mapPq, VqLq, Qf, rhsQf, rhsQ, (hipStream_t)stream);
}

int wadges_update(const wadges_plan *p, int K, const double *Jq, const double *VqPq,
                  const double *VfPq, double fa, double fb, double fdt, double *rhsQ,
                  double *resQ, double *Q, double *Qf, void *stream) {
  if (!p || K < 0) return WADGES_EINVAL;
  if (K > 0 && (!Jq || !VqPq || !VfPq || !rhsQ || !resQ || !Q || !Qf)) return WADGES_EINVAL;
  return p->ops.update(p, K, Jq, VqPq, VfPq, fa, fb, fdt, rhsQ, resQ, Q, Qf, (hipStream_t)stream);
}

int wadges_diagnostics(const wadges_plan *p, int K, const double *wq, const double *Jq,
                       const double *Q, const double *Qexact, double *aux, double *totals,
                       void *stream) {
  if (!p || K < 0 || !aux || !totals) return WADGES_EINVAL;
  if (K > 0 && (!wq || !Jq || !Q)) return WADGES_EINVAL;
  return p->ops.diag(p, K, wq, Jq, Q, Qexact, aux, totals, (hipStream_t)stream);
}

int wadges_diagnostics_noalloc(const wadges_plan *p, int K, const double *wq, const double *Jq,
                               const double *Q, const double *Qexact, double *totals,
                               void *stream) {
  if (!p || K < 0 || !totals) return WADGES_EINVAL;
  if (K > 0 && (!wq || !Jq || !Q)) return WADGES_EINVAL;
  if (K > WADGES_NOALLOC_MAX_K) return WADGES_ETOOLARGE;
  double *aux = p->scratch_addr;
  if (!aux) {
    int dev = -1;
    int rc = hip_check(hipGetDevice(&dev));
    if (rc) return rc;
    if (dev != p->device) return WADGES_EINVAL;
    rc = scratch_address(dev, &aux);
    if (rc) return rc;
  }
  // every entry of aux[0, 4K) is written by diag_kernel before the reduction
  // reads it, so the array needs no clearing between calls
  return p->ops.diag(p, K, wq, Jq, Q, Qexact, aux, totals, (hipStream_t)stream);
}

int wadges_rhs_vjp(const wadges_plan *p, int K, const double *vgeo, const double *vfgeo,
                   const double *fgeo, const uint32_t *mapPq, const double *nrJ, const double *nsJ,
                   const double *ntJ, const double *Drq, const double *Dsq, const double *Dtq,
                   const double *VqLq, const double *VfPq, const double *rhsQ, const double *Qf,
                   const double *rhsQ_bar, double *sbar, double *rhsQ_state_bar,
                   double *Qf_bar, void *stream) {
  if (!p || K < 0) return WADGES_EINVAL;
  if (K > 0 && (!vgeo || !vfgeo || !fgeo || !mapPq || !nrJ || !nsJ || !Drq || !Dsq || !VqLq ||
                !VfPq || !rhsQ || !Qf || !rhsQ_bar || !sbar || !rhsQ_state_bar || !Qf_bar))
    return WADGES_EINVAL;
  if (K > 0 && p->dim == 3 && (!ntJ || !Dtq)) return WADGES_EINVAL;
  return p->ops.rhs_vjp(p, K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, Drq, Dsq, Dtq, VqLq, VfPq,
                        rhsQ, Qf, rhsQ_bar, sbar, rhsQ_state_bar, Qf_bar, (hipStream_t)stream);
}

int wadges_update_vjp(const wadges_plan *p, int K, const double *Jq, const double *VqPq,
                      const double *VfPq, double fa, double fb, double fdt, const double *Q,
                      double *Q_bar, double *resQ_bar, const double *rhsQ_bar_out,
                      const double *Qf_bar_out, double *rhsQ_bar, void *stream) {
  if (!p || K < 0) return WADGES_EINVAL;
  if (K > 0 && (!Jq || !VqPq || !VfPq || !Q || !Q_bar || !resQ_bar || !rhsQ_bar_out ||
                !Qf_bar_out || !rhsQ_bar))
    return WADGES_EINVAL;
  return p->ops.update_vjp(p, K, Jq, VqPq, VfPq, fa, fb, fdt, Q, Q_bar, resQ_bar, rhsQ_bar_out,
                           Qf_bar_out, rhsQ_bar, (hipStream_t)stream);
}

int wadges_project_vjp(const wadges_plan *p, int K, const double *VqPq, const double *VfPq,
                       const double *Q, const double *rhsQ_bar, const double *Qf_bar,
                       double *Q_bar, void *stream) {
  if (!p || K < 0) return WADGES_EINVAL;
  if (K > 0 && (!VqPq || !VfPq || !Q || !rhsQ_bar || !Qf_bar || !Q_bar)) return WADGES_EINVAL;
  return p->ops.proj_vjp(p, K, VqPq, VfPq, Q, rhsQ_bar, Qf_bar, Q_bar, (hipStream_t)stream);
}

int wadges_stage(const wadges_plan *p, int K, const double *vgeo, const double *vfgeo,
                 const double *fgeo, const uint32_t *mapPq, const double *nrJ, const double *nsJ,
                 const double *ntJ, const double *Drq, const double *Dsq, const double *Dtq,
                 const double *VqLq, const double *VqPq, const double *VfPq, const double *Jq,
                 double fa, double fb, double fdt, const double *Qf_in, double *Qf_out,
                 double *rhsQ, double *resQ, double *Q, void *stream) {
  if (!p || K < 0) return WADGES_EINVAL;
  if (K > 0 && (!vgeo || !vfgeo || !fgeo || !mapPq || !nrJ || !nsJ || !Drq || !Dsq || !VqLq ||
                !VqPq || !VfPq || !Jq || !Qf_in || !Qf_out || !rhsQ || !resQ || !Q))
    return WADGES_EINVAL;
  if (K > 0 && p->dim == 3 && (!ntJ || !Dtq)) return WADGES_EINVAL;
  if (K > 0 && Qf_in == Qf_out) return WADGES_EINVAL;  // neighbours read Qf_in concurrently
  return p->ops.stage(p, K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, Drq, Dsq, Dtq, VqLq, VqPq,
                      VfPq, Jq, fa, fb, fdt, Qf_in, Qf_out, rhsQ, rhsQ, resQ, resQ, Q, Q,
                      (hipStream_t)stream);
}

int wadges_stage_record(const wadges_plan *p, int K, const double *vgeo, const double *vfgeo,
                        const double *fgeo, const uint32_t *mapPq, const double *nrJ,
                        const double *nsJ, const double *ntJ, const double *Drq, const double *Dsq,
                        const double *Dtq, const double *VqLq, const double *VqPq,
                        const double *VfPq, const double *Jq, double fa, double fb, double fdt,
                        const double *Qf_in, double *Qf_out, const double *rhsQ_in,
                        double *rhsQ_out, const double *resQ_in, double *resQ_out,
                        const double *Q_in, double *Q_out, void *stream) {
  if (!p || K < 0) return WADGES_EINVAL;
  if (K > 0 && (!vgeo || !vfgeo || !fgeo || !mapPq || !nrJ || !nsJ || !Drq || !Dsq || !VqLq ||
                !VqPq || !VfPq || !Jq || !Qf_in || !Qf_out || !rhsQ_in || !rhsQ_out ||
                !resQ_in || !resQ_out || !Q_in || !Q_out))
    return WADGES_EINVAL;
  if (K > 0 && p->dim == 3 && (!ntJ || !Dtq)) return WADGES_EINVAL;
  if (K > 0 && Qf_in == Qf_out) return WADGES_EINVAL;
  return p->ops.stage(p, K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, Drq, Dsq, Dtq, VqLq, VqPq,
                      VfPq, Jq, fa, fb, fdt, Qf_in, Qf_out, rhsQ_in, rhsQ_out, resQ_in, resQ_out,
                      Q_in, Q_out, (hipStream_t)stream);
}

int wadges_stage_vjp(const wadges_plan *p, int K, const double *vgeo, const double *vfgeo,
                     const double *fgeo, const uint32_t *mapPq, const double *nrJ,
                     const double *nsJ, const double *ntJ, const double *Drq, const double *Dsq,
                     const double *Dtq, const double *VqLq, const double *VqPq, const double *VfPq,
                     const double *Jq, double fa, double fb, double fdt, const double *rhsQ_in,
                     const double *Qf_in, const double *Q_out, double *Q_bar, double *resQ_bar,
                     const double *rhsQ_bar_out, const double *Qf_bar_out, double *R_bar,
                     double *sbar, double *rhsQ_bar_in, double *Qf_bar_in, void *stream) {
  if (!p || K < 0) return WADGES_EINVAL;
  if (K > 0 && (!vgeo || !vfgeo || !fgeo || !mapPq || !nrJ || !nsJ || !Drq || !Dsq || !VqLq ||
                !VqPq || !VfPq || !Jq || !rhsQ_in || !Qf_in || !Q_out || !Q_bar || !resQ_bar ||
                !rhsQ_bar_out || !Qf_bar_out || !R_bar || !sbar || !rhsQ_bar_in || !Qf_bar_in))
    return WADGES_EINVAL;
  if (K > 0 && p->dim == 3 && (!ntJ || !Dtq)) return WADGES_EINVAL;
  return p->ops.stage_vjp(p, K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, Drq, Dsq, Dtq, VqLq,
                          VqPq, VfPq, Jq, fa, fb, fdt, rhsQ_in, Qf_in, Q_out, Q_bar, resQ_bar,
                          rhsQ_bar_out, Qf_bar_out, R_bar, sbar, rhsQ_bar_in, Qf_bar_in,
                          (hipStream_t)stream);
}

}  // extern "C"
