// wadges_batch.hpp -- ensemble (batched) LSRK stage: many independent meshes of
// one dimension, uneven K and N, advanced by one launch per stage.
//
// Included by wadges.hip at file scope, after the instance table.
//
// Device side.  The work of a stage is a list of items (problem, first
// element); an item is tiles_per_item tiles of S::TPW elements, the
// elements one 64-wide workgroup of the fused stage kernel holds.  A workgroup
// walks items in a static stride or takes them from an atomic counter, and for
// each runs stage_body of the item's (Nq, Nfq) instance -- the same device code
// as wadges_stage, so every element's arithmetic is the single-problem one and
// results do not depend on the schedule.  Dynamic LDS is the largest per-tile
// need of the instances present.  The work counter is reset by the last
// workgroup to finish, so launches need no memset (and can be graph-captured).
//
// Host side.  wadges_batch owns the device copies of the problem descriptors,
// the per-problem operator tables (prepared once), the item list of the current
// launch shape, the work counter, a capture stream and two HIP graphs of one
// LSRK45 step (one per starting face-trace buffer: five stages flip it).

namespace wadges {

// Problem descriptor as the kernel reads it.
struct BatchGroup {
  StageArgs a;      // fa, fb, Qf_in, Qf_out unused: they come per launch
  double *qf[2];
  int slot;         // index of the (Nq, Nfq) instance within its dimension
  int pad;
};

// Descriptors are read through the constant address space (see ConstArgs).
typedef __attribute__((address_space(4))) const BatchGroup ConstGroup;

template <class... Ss> struct ShapeList {};
using Family2 = ShapeList<Shape<2, 3, 2>, Shape<2, 6, 3>, Shape<2, 12, 4>, Shape<2, 16, 5>,
                          Shape<2, 25, 6>>;
using Family3 = ShapeList<Shape<3, 4, 3>, Shape<3, 11, 6>, Shape<3, 23, 12>, Shape<3, 44, 16>>;

template <class S>
__device__ __forceinline__ void batch_item(const ConstArgs *A, double fa, double fb,
                                           const double *Qf_in, double *Qf_out, int k0, int tiles,
                                           double (*s_nref)[S::DIM], double *smem) {
  constexpr int TPW = S::TPW;
  for (int t = 0; t < tiles; ++t) {
    const int kfirst = k0 + t * TPW;
    if (kfirst >= A->K) break;          // uniform over the workgroup
    // An opaque copy of the lane index per tile keeps the compiler from hoisting
    // lane-dependent schedule indices out of the loops (they would stay live
    // across them and multiply the register footprint).
    int lane;
    asm volatile("v_mov_b32 %0, %1" : "=v"(lane) : "v"((int)threadIdx.x));
    stage_body<S>(A, fa, fb, Qf_in, Qf_out, kfirst, lane, s_nref, smem);
    __syncthreads();                    // LDS is reused by the next tile
  }
}

template <int DIM, class... Ss>
__device__ __forceinline__ void batch_dispatch(ShapeList<Ss...>, int slot, const ConstArgs *A,
                                               double fa, double fb, const double *Qf_in,
                                               double *Qf_out, int k0, int tiles,
                                               double (*s_nref)[DIM], double *smem) {
  int idx = 0;
  ((slot == idx++ ? batch_item<Ss>(A, fa, fb, Qf_in, Qf_out, k0, tiles, s_nref, smem) : void()),
   ...);
}

// Persistent form: a workgroup loops over items.  Loop-invariant constants of
// the transcendental functions would be hoisted out of the loop and held in
// registers; the occupancy hint makes the register allocator rematerialise them
// instead (2D: 128 VGPRs, four waves per SIMD; 3D: two waves per SIMD, which
// the LDS of the 3D shapes allows at most, leaving room for the two rows per
// lane of 3D N = 3).
template <int DIM> constexpr int batch_waves() { return DIM == 2 ? 4 : 2; }

template <int DIM, class Fam>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(batch_waves<DIM>())))
batch_stage_persistent(const BatchGroup *__restrict__ groups, const int2 *__restrict__ items,
                       int nitems, int tiles, double fa, double fb, int cur, int *queue) {
  extern __shared__ double smem[];
  __shared__ double s_nref[DIM + 1][DIM];
  __shared__ int s_next;
  int it = blockIdx.x;
  while (it < nitems) {
    const int2 w = items[it];
    const ConstGroup *G = (const ConstGroup *)groups + w.x;
    batch_dispatch<DIM>(Fam{}, G->slot, &G->a, fa, fb, G->qf[cur], G->qf[cur ^ 1], w.y, tiles,
                        s_nref, smem);
    if (queue) {
      if (threadIdx.x == 0) s_next = (int)gridDim.x + atomicAdd(queue, 1);
      __syncthreads();
      it = __builtin_amdgcn_readfirstlane(s_next);
      __syncthreads();
    } else {
      it += gridDim.x;
    }
  }
  if (queue && threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(queue + 1, 1) == (int)gridDim.x - 1) {
      __threadfence();
      atomicExch(queue, 0);
      atomicExch(queue + 1, 0);
    }
  }
}

// One tile of a workgroup of the flat launch: stage_body for every wave, also
// for a padding tile past the problem's last element (all lanes idle), so that
// all waves of a workgroup reach the same barriers.
template <class S>
__device__ __forceinline__ void flat_tile(const ConstArgs *A, double fa, double fb,
                                          const double *Qf_in, double *Qf_out, int kfirst,
                                          int lane, double (*s_nref)[S::DIM], double *smem) {
  int l;
  asm volatile("v_mov_b32 %0, %1" : "=v"(l) : "v"(lane));
  stage_body<S>(A, fa, fb, Qf_in, Qf_out, kfirst, l, s_nref, smem);
}

template <int DIM, class... Ss>
__device__ __forceinline__ void flat_dispatch(ShapeList<Ss...>, int slot, const ConstArgs *A,
                                              double fa, double fb, const double *Qf_in,
                                              double *Qf_out, int k0, int lane,
                                              double (*s_nref)[DIM], double *smem) {
  int idx = 0;
  ((slot == idx++ ? flat_tile<Ss>(A, fa, fb, Qf_in, Qf_out, k0, lane, s_nref, smem) : void()),
   ...);
}

// One tile per wave, W waves per workgroup, no loop: the register use of the
// single-problem stage kernels (the largest over the family).  The W tiles of
// a workgroup belong to one problem (the item list pads each problem to a
// multiple of W tiles), so its waves run the same shape and meet the same
// barriers; each wave has its own LDS region of smem_wave doubles.  2D: four
// waves per SIMD declared; 3D: two, which the LDS of 3D N = 3 allows at most
// (unconstrained, the matrix-core accumulators of the stage body went to extra
// registers and left one wave per SIMD).
template <int DIM> constexpr int flat_waves() { return DIM == 2 ? 4 : 2; }

template <int DIM, class Fam, int W>
__global__ void __launch_bounds__(WAVE * W) __attribute__((amdgpu_waves_per_eu(flat_waves<DIM>())))
batch_stage_flat(const BatchGroup *__restrict__ groups, const int2 *__restrict__ items,
                 double fa, double fb, int cur, int smem_wave) {
  extern __shared__ double smem[];
  __shared__ double s_nref[W][DIM + 1][DIM];
  const int wave = W > 1 ? __builtin_amdgcn_readfirstlane((int)threadIdx.x / WAVE) : 0;
  const int lane = W > 1 ? (int)threadIdx.x - wave * WAVE : (int)threadIdx.x;
  const int2 w = items[blockIdx.x * W + wave];
  const ConstGroup *G = (const ConstGroup *)groups + w.x;
  flat_dispatch<DIM>(Fam{}, G->slot, &G->a, fa, fb, G->qf[cur], G->qf[cur ^ 1], w.y, lane,
                     s_nref[wave], smem + (size_t)wave * smem_wave);
}

// LSRK45 coefficients (Carpenter & Kennedy 1994), as tabulated in euler.c:2779-2790.
constexpr double kRK4A[5] = {0.0, -567301805773.0 / 1357537059087.0,
                             -2404267990393.0 / 2016746695238.0,
                             -3550918686646.0 / 2091501179385.0,
                             -1275806237668.0 / 842570457699.0};
constexpr double kRK4B[5] = {1432997174477.0 / 9575080441755.0,
                             5161836677717.0 / 13612068292357.0,
                             1720146321549.0 / 2090206949498.0,
                             3134564353537.0 / 4481467310338.0,
                             2277821191437.0 / 14882151754819.0};

}  // namespace wadges

struct wadges_batch {
  int dim = 0, ngroups = 0, device = 0, cus = 0;
  double gamma = 1.4, tau = 1.0;
  std::vector<wadges_group> g;           // caller descriptors
  std::vector<int> slot, tpw;            // per problem
  std::vector<double> cost;              // per element, relative
  wadges::BatchGroup *d_groups = nullptr;
  double *ws = nullptr;                  // operator tables of all problems
  int *queue = nullptr;                  // work counter + finished-workgroup counter
  int2 *d_items = nullptr;
  int nitems = 0, items_cap = 0;
  size_t smem = 0;                       // dynamic LDS bytes per workgroup
  wadges_batch_config cfg{1, 0, 0, 0, 1};
  int cur = 0;
  hipStream_t cap = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
};

namespace wadges {

inline const void *batch_persistent_ptr(int dim) {
  return dim == 2 ? (const void *)&batch_stage_persistent<2, Family2>
                  : (const void *)&batch_stage_persistent<3, Family3>;
}

inline void batch_drop_graphs(wadges_batch *b) {
  for (int p = 0; p < 2; ++p) {
    if (b->exec[p]) (void)hipGraphExecDestroy(b->exec[p]);
    if (b->graph[p]) (void)hipGraphDestroy(b->graph[p]);
    b->exec[p] = nullptr;
    b->graph[p] = nullptr;
  }
}

// Item list of a launch shape: (problem, first element) per item, in problem
// order or costliest first (stable).  The flat launch takes items in
// workgroups of waves_per_group tiles of one problem, each problem padded with
// tiles past its last element to a multiple of that.  Uploaded synchronously:
// launch shapes change outside the stepping loop.
inline int batch_build_items(wadges_batch *b, wadges_batch_config c) {
  if (c.grid == 0) c.tiles_per_item = 1;   // the flat kernel takes one tile per item
  if (c.grid == 0) c.dynamic = 0;
  if (c.grid > 0) c.waves_per_group = 1;   // the persistent kernel runs one-wave workgroups
  const int W = c.waves_per_group;
  std::vector<std::vector<int2>> units;    // items kept together (a flat workgroup's tiles)
  std::vector<double> ucost;
  for (int q = 0; q < b->ngroups; ++q) {
    const int per = c.tiles_per_item * b->tpw[q];
    std::vector<int2> unit;
    double cost = 0.0;
    for (int k0 = 0; k0 < b->g[q].K; k0 += per) {
      unit.push_back(make_int2(q, k0));
      cost += b->cost[q] * std::min(per, b->g[q].K - k0);
      if ((int)unit.size() == W) {
        units.push_back(unit);
        ucost.push_back(cost);
        unit.clear();
        cost = 0.0;
      }
    }
    if (!unit.empty()) {
      while ((int)unit.size() < W) unit.push_back(make_int2(q, b->g[q].K));   // idle tiles
      units.push_back(unit);
      ucost.push_back(cost);
    }
  }
  std::vector<size_t> idx(units.size());
  for (size_t n = 0; n < idx.size(); ++n) idx[n] = n;
  if (c.order == 1)
    std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) { return ucost[x] > ucost[y]; });
  std::vector<int2> items;
  for (size_t n : idx) items.insert(items.end(), units[n].begin(), units[n].end());
  const int n = (int)items.size();
  if (n > b->items_cap) {
    if (b->d_items) {
      int rc = hip_check(hipFree(b->d_items));
      b->d_items = nullptr;
      b->items_cap = 0;
      if (rc) return rc;
    }
    hipError_t e = hipMalloc(&b->d_items, sizeof(int2) * (size_t)n);
    if (e != hipSuccess) {
      g_last_hip = (int)e;
      return WADGES_ENOMEM;
    }
    b->items_cap = n;
  }
  if (n > 0) {
    int rc = hip_check(hipMemcpy(b->d_items, items.data(), sizeof(int2) * (size_t)n,
                                 hipMemcpyHostToDevice));
    if (rc) return rc;
  }
  b->nitems = n;
  b->cfg = c;
  batch_drop_graphs(b);
  return WADGES_OK;
}

template <int DIM, class Fam>
inline void launch_flat(const wadges_batch *b, int W, double fa, double fb, int cur,
                        hipStream_t st) {
  const int blocks = b->nitems / W;
  const int sw = (int)(b->smem / sizeof(double));
  const size_t lds = b->smem * (size_t)W;
  if (W == 4)
    batch_stage_flat<DIM, Fam, 4><<<blocks, WAVE * 4, lds, st>>>(b->d_groups, b->d_items, fa, fb,
                                                                cur, sw);
  else if (W == 2)
    batch_stage_flat<DIM, Fam, 2><<<blocks, WAVE * 2, lds, st>>>(b->d_groups, b->d_items, fa, fb,
                                                                cur, sw);
  else
    batch_stage_flat<DIM, Fam, 1><<<blocks, WAVE, lds, st>>>(b->d_groups, b->d_items, fa, fb, cur,
                                                            sw);
}

inline int batch_launch(wadges_batch *b, double fa, double fb, int cur, hipStream_t st) {
  if (b->nitems == 0) return WADGES_OK;
  const int tiles = b->cfg.tiles_per_item;
  if (b->cfg.grid == 0) {
    if (b->dim == 2)
      launch_flat<2, Family2>(b, b->cfg.waves_per_group, fa, fb, cur, st);
    else
      launch_flat<3, Family3>(b, b->cfg.waves_per_group, fa, fb, cur, st);
    return hip_check(hipGetLastError());
  }
  const int grid = std::min(b->cfg.grid, b->nitems);
  int *queue = (b->cfg.dynamic && grid < b->nitems) ? b->queue : nullptr;
  if (b->dim == 2)
    batch_stage_persistent<2, Family2><<<grid, WAVE, b->smem, st>>>(
        b->d_groups, b->d_items, b->nitems, tiles, fa, fb, cur, queue);
  else
    batch_stage_persistent<3, Family3><<<grid, WAVE, b->smem, st>>>(
        b->d_groups, b->d_items, b->nitems, tiles, fa, fb, cur, queue);
  return hip_check(hipGetLastError());
}

// Relative per-element cost of a stage: two-point flux evaluations per element
// (symmetric volume pairs + volume-face pairs), the dominant term.
inline double batch_elem_cost(int Nq, int Nf) {
  return 0.5 * Nq * (Nq + 1.0) + (double)Nq * Nf + 4.0 * Nq;
}

inline std::string batch_signature(const wadges_batch *b) {
  std::vector<std::pair<int, int>> v;
  for (const auto &q : b->g) v.push_back({q.Nq, q.K});
  std::sort(v.begin(), v.end());
  std::string s = "d" + std::to_string(b->dim) + ":cu" + std::to_string(b->cus);
  for (const auto &p : v) s += ":" + std::to_string(p.first) + "x" + std::to_string(p.second);
  return s;
}

inline std::map<std::string, wadges_batch_config> &tune_cache() {
  static std::map<std::string, wadges_batch_config> m;
  return m;
}

inline bool tune_file_lookup(const std::string &sig, wadges_batch_config *c) {
  const char *path = std::getenv("WADGES_TUNE_CACHE");
  if (!path || !*path) return false;
  std::ifstream f(path);
  std::string key;
  wadges_batch_config t;
  bool found = false;
  while (f >> key >> t.tiles_per_item >> t.grid >> t.dynamic >> t.order >> t.waves_per_group)
    if (key == sig) {
      *c = t;
      found = true;
    }
  return found;
}

inline void tune_file_store(const std::string &sig, const wadges_batch_config &c) {
  const char *path = std::getenv("WADGES_TUNE_CACHE");
  if (!path || !*path) return;
  std::ofstream f(path, std::ios::app);
  f << sig << ' ' << c.tiles_per_item << ' ' << c.grid << ' ' << c.dynamic << ' ' << c.order
    << ' ' << c.waves_per_group << '\n';
}

}  // namespace wadges

// ================================================================== C ABI

extern "C" {

int wadges_batch_create(wadges_batch **out, int dim, int ngroups, const wadges_group *groups,
                        double gamma, double tau, void *stream) {
  using namespace wadges;
  if (!out) return WADGES_EINVAL;
  *out = nullptr;
  if ((dim != 2 && dim != 3) || ngroups < 0 || (ngroups > 0 && !groups) || !(gamma > 1.0) ||
      !(tau >= 0.0))
    return WADGES_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  std::unique_ptr<wadges_batch> b(new (std::nothrow) wadges_batch);
  if (!b) return WADGES_ENOMEM;
  b->dim = dim;
  b->ngroups = ngroups;
  b->gamma = gamma;
  b->tau = tau;
  b->g.assign(groups, groups + ngroups);
  std::vector<const Instance *> inst(ngroups, nullptr);
  size_t ws_total = 0;
  for (int q = 0; q < ngroups; ++q) {
    const wadges_group &G = groups[q];
    int slot = 0;
    for (int n = 0; n < kNumInstances; ++n) {
      if (kInstances[n].dim != dim) continue;
      if (kInstances[n].Nq == G.Nq && kInstances[n].Nfq == G.Nfq) {
        inst[q] = &kInstances[n];
        break;
      }
      ++slot;
    }
    if (!inst[q]) return WADGES_EUNSUPPORTED;
    if (G.K < 0 || !std::isfinite(G.dt)) return WADGES_EINVAL;
    if (G.K > 0 && (!G.vgeo || !G.vfgeo || !G.fgeo || !G.mapPq || !G.nrJ || !G.nsJ || !G.Drq ||
                    !G.Dsq || !G.VqLq || !G.VqPq || !G.VfPq || !G.Jq || !G.Q || !G.rhsQ ||
                    !G.resQ || !G.Qf[0] || !G.Qf[1] || G.Qf[0] == G.Qf[1]))
      return WADGES_EINVAL;
    if (G.K > 0 && dim == 3 && (!G.ntJ || !G.Dtq)) return WADGES_EINVAL;
    b->slot.push_back(slot);
    b->tpw.push_back(inst[q]->ops.stage_tpw);
    b->cost.push_back(batch_elem_cost(G.Nq, G.Nfq * (dim + 1)));
    b->smem = std::max(b->smem, (size_t)inst[q]->ops.stage_smem);
    ws_total += (size_t)inst[q]->ops.ws_doubles;
  }
  int rc = hip_check(hipGetDevice(&b->device));
  if (!rc) rc = hip_check(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount,
                                                b->device));
  if (rc) return rc;
  auto fail = [&](int code) {
    wadges_batch_destroy(b.release());
    return code;
  };
  if (hipMalloc(&b->queue, 2 * sizeof(int)) != hipSuccess) return fail(WADGES_ENOMEM);
  if ((rc = hip_check(hipMemset(b->queue, 0, 2 * sizeof(int))))) return fail(rc);
  if ((rc = hip_check(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking)))) return fail(rc);
  if (ngroups > 0) {
    if (hipMalloc(&b->ws, sizeof(double) * std::max<size_t>(ws_total, 1)) != hipSuccess)
      return fail(WADGES_ENOMEM);
    if (hipMalloc(&b->d_groups, sizeof(BatchGroup) * (size_t)ngroups) != hipSuccess)
      return fail(WADGES_ENOMEM);
    std::vector<BatchGroup> hg(ngroups);
    const double gm1 = gamma - 1.0;
    size_t off = 0;
    for (int q = 0; q < ngroups; ++q) {
      const wadges_group &G = groups[q];
      double *ws = b->ws + off;
      off += (size_t)inst[q]->ops.ws_doubles;
      if (G.K > 0 && (rc = inst[q]->ops.prep(G.Drq, G.Dsq, G.Dtq, G.VqLq, G.VfPq, ws, st)))
        return fail(rc);
      hg[q].a = StageArgs{G.K,    G.vgeo, G.vfgeo, G.fgeo,  G.mapPq, G.nrJ,  G.nsJ,
                          G.ntJ,  ws,     G.VqLq,  G.VqPq,  G.VfPq,  G.Jq,   0.0,
                          0.0,    G.dt,   nullptr, nullptr, G.rhsQ,  G.resQ, G.Q,
                          G.rhsQ, G.resQ, G.Q,
                          gamma,  gm1,    1.0 / (2.0 * gm1), 0.25 * tau, std::log(gm1)};
      hg[q].qf[0] = G.Qf[0];
      hg[q].qf[1] = G.Qf[1];
      hg[q].slot = b->slot[q];
      hg[q].pad = 0;
    }
    if ((rc = hip_check(hipMemcpy(b->d_groups, hg.data(), sizeof(BatchGroup) * (size_t)ngroups,
                                  hipMemcpyHostToDevice))))
      return fail(rc);
  }
  if ((rc = batch_build_items(b.get(), b->cfg))) return fail(rc);
  if ((rc = hip_check(hipStreamSynchronize(st)))) return fail(rc);
  *out = b.release();
  return WADGES_OK;
}

int wadges_batch_destroy(wadges_batch *b) {
  using namespace wadges;
  if (!b) return WADGES_OK;
  int rc = WADGES_OK;
  batch_drop_graphs(b);
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess && rc == WADGES_OK) rc = hip_check(e);
  };
  if (b->cap) chk(hipStreamDestroy(b->cap));
  if (b->d_items) chk(hipFree(b->d_items));
  if (b->d_groups) chk(hipFree(b->d_groups));
  if (b->ws) chk(hipFree(b->ws));
  if (b->queue) chk(hipFree(b->queue));
  delete b;
  return rc;
}

int wadges_batch_stage(wadges_batch *b, double fa, double fb, void *stream) {
  using namespace wadges;
  if (!b) return WADGES_EINVAL;
  int rc = batch_launch(b, fa, fb, b->cur, (hipStream_t)stream);
  if (!rc) b->cur ^= 1;
  return rc;
}

int wadges_batch_run(wadges_batch *b, int nsteps, int use_graph, void *stream) {
  using namespace wadges;
  if (!b || nsteps < 0) return WADGES_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  for (int n = 0; n < nsteps; ++n) {
    if (!use_graph) {
      for (int s = 0; s < 5; ++s) {
        int rc = batch_launch(b, kRK4A[s], kRK4B[s], b->cur, st);
        if (rc) return rc;
        b->cur ^= 1;
      }
      continue;
    }
    const int p = b->cur;
    if (!b->exec[p]) {
      int rc = hip_check(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
      if (rc) return rc;
      int c = p;
      for (int s = 0; s < 5 && !rc; ++s) {
        rc = batch_launch(b, kRK4A[s], kRK4B[s], c, b->cap);
        c ^= 1;
      }
      hipGraph_t gr = nullptr;
      const int rc_end = hip_check(hipStreamEndCapture(b->cap, &gr));
      if (rc || rc_end) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc ? rc : rc_end;
      }
      b->graph[p] = gr;
      rc = hip_check(hipGraphInstantiate(&b->exec[p], gr, nullptr, nullptr, 0));
      if (rc) {
        b->exec[p] = nullptr;
        return rc;
      }
    }
    int rc = hip_check(hipGraphLaunch(b->exec[p], st));
    if (rc) return rc;
    b->cur ^= 1;  // five stages
  }
  return WADGES_OK;
}

int wadges_batch_current(const wadges_batch *b) { return b ? b->cur : WADGES_EINVAL; }

int wadges_batch_get_config(const wadges_batch *b, wadges_batch_config *cfg) {
  if (!b || !cfg) return WADGES_EINVAL;
  *cfg = b->cfg;
  return WADGES_OK;
}

int wadges_batch_set_config(wadges_batch *b, const wadges_batch_config *cfg) {
  using namespace wadges;
  if (!b || !cfg) return WADGES_EINVAL;
  if (cfg->tiles_per_item < 1 || cfg->tiles_per_item > 64 || cfg->grid < 0 ||
      (cfg->dynamic != 0 && cfg->dynamic != 1) || (cfg->order != 0 && cfg->order != 1) ||
      (cfg->waves_per_group != 1 && cfg->waves_per_group != 2 && cfg->waves_per_group != 4))
    return WADGES_EINVAL;
  // the flat launch's LDS is waves_per_group tiles' worth
  if (cfg->grid == 0 && b->smem * (size_t)cfg->waves_per_group > 65536) return WADGES_EINVAL;
  int rc = hip_check(hipDeviceSynchronize());  // the item list may be in use
  if (rc) return rc;
  return batch_build_items(b, *cfg);
}

int wadges_batch_tune(wadges_batch *b, int reps, double *ms, void *stream) {
  using namespace wadges;
  if (!b) return WADGES_EINVAL;
  if (ms) *ms = -1.0;
  hipStream_t st = (hipStream_t)stream;
  if (reps <= 0) reps = 5;
  const std::string sig = batch_signature(b);
  wadges_batch_config found{};
  auto &cache = tune_cache();
  auto hit = cache.find(sig);
  if (hit != cache.end()) return wadges_batch_set_config(b, &hit->second);
  if (tune_file_lookup(sig, &found)) {
    cache[sig] = found;
    return wadges_batch_set_config(b, &found);
  }
  int rc = hip_check(hipDeviceSynchronize());
  if (rc) return rc;
  // Save the state (Q, rhsQ, resQ and the current face traces of every problem):
  // candidates are timed on real stages.
  const int nf = b->dim + 2;
  size_t total = 0;
  for (const auto &G : b->g)
    total += (size_t)G.K * nf * (3 * (size_t)G.Nq + (size_t)G.Nfq * (b->dim + 1));
  double *save = nullptr;
  if (total > 0 && hipMalloc(&save, sizeof(double) * total) != hipSuccess) return WADGES_ENOMEM;
  const int cur0 = b->cur;
  const wadges_batch_config cfg0 = b->cfg;
  auto copy_state = [&](bool to_save) -> int {
    size_t off = 0;
    for (const auto &G : b->g) {
      const size_t nq = (size_t)G.K * nf * G.Nq, nfc = (size_t)G.K * nf * G.Nfq * (b->dim + 1);
      double *arr[4] = {G.Q, G.rhsQ, G.resQ, G.Qf[cur0]};
      const size_t len[4] = {nq, nq, nq, nfc};
      for (int a = 0; a < 4; ++a) {
        if (len[a] == 0) continue;
        double *dst = to_save ? save + off : arr[a];
        const double *src = to_save ? arr[a] : save + off;
        int r = hip_check(hipMemcpyAsync(dst, src, sizeof(double) * len[a],
                                         hipMemcpyDeviceToDevice, st));
        if (r) return r;
        off += len[a];
      }
    }
    return WADGES_OK;
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double best_ms = 1e300;
  wadges_batch_config best = cfg0;
  rc = copy_state(true);
  if (!rc) rc = hip_check(hipEventCreate(&e0));
  if (!rc) rc = hip_check(hipEventCreate(&e1));
  int occ = 1;
  if (!rc)
    rc = hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &occ, batch_persistent_ptr(b->dim), WAVE, b->smem));
  const int resident = std::max(1, occ) * std::max(1, b->cus);
  // candidates: the flat launch (one tile per wave) with workgroups of 1, 2 or 4
  // waves (as LDS allows) in both orders, and persistent launches of one or two
  // resident waves' worth of one-wave workgroups per CU, static or dynamic,
  // over items of 1, 2 or 4 tiles, in both orders
  std::vector<wadges_batch_config> cand;
  for (int W : {1, 2, 4})
    if (W == 1 || b->smem * (size_t)W <= 65536)
      for (int order = 0; order < 2; ++order) cand.push_back({1, 0, 0, order, W});
  for (int tiles : {1, 2, 4})
    for (int m : {1, 2})
      for (int dyn = 0; dyn < 2; ++dyn)
        for (int order = 0; order < 2; ++order)
          cand.push_back({tiles, m * resident, dyn, order, 1});
  if (!rc) {
    for (const auto &c : cand) {
        if ((rc = batch_build_items(b, c))) break;
        if ((rc = batch_launch(b, kRK4A[0], kRK4B[0], b->cur, st))) break;  // warm-up
        b->cur ^= 1;
        if ((rc = hip_check(hipEventRecord(e0, st)))) break;
        for (int r = 0; r < reps && !rc; ++r) {
          rc = batch_launch(b, kRK4A[(r + 1) % 5], kRK4B[(r + 1) % 5], b->cur, st);
          b->cur ^= 1;
        }
        if (rc || (rc = hip_check(hipEventRecord(e1, st)))) break;
        if ((rc = hip_check(hipEventSynchronize(e1)))) break;
        float t = 0.f;
        if ((rc = hip_check(hipEventElapsedTime(&t, e0, e1)))) break;
        if (t / reps < best_ms) {
          best_ms = t / reps;
          best = c;
        }
    }
  }
  // restore the state whatever happened
  b->cur = cur0;
  int rc2 = copy_state(false);
  if (!rc2) rc2 = hip_check(hipStreamSynchronize(st));
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (save) (void)hipFree(save);
  if (!rc) rc = rc2;
  const int rc3 = batch_build_items(b, rc ? cfg0 : best);
  if (rc) return rc;
  if (rc3) return rc3;
  cache[sig] = best;
  tune_file_store(sig, best);
  if (ms) *ms = best_ms;
  return WADGES_OK;
}

}  // extern "C"
