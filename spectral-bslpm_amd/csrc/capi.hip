// capi.hip -- the C ABI (include/sbslpm.h): context, workspace, entry points
// and the bve_back_rk4_step orchestration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <initializer_list>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/sbslpm.h"
#include "device.hpp"
#include "irreg_tree.hpp"
#include "tree.hpp"

using namespace sbslpm;

// Per-evaluation state of a BVE step for `members` stacked members: particle
// arrays [members][Pl][m2], proxy arrays [members][P][n], records per member.
struct Work {
  int members = 0;
  // precision and multi-RHS columns the record buffers were sized for (an
  // ensemble workspace holds only the record kind of its precision)
  int precision = -1;
  bool mrhs = false;
  SrcRec64* rec64 = nullptr;
  SrcRec32* rec32 = nullptr;
  SrcRecMfma* recmf = nullptr;
  float* qcol = nullptr;  // multi-RHS charge columns (ensembles with SBSLPM_FLAG_MULTI_RHS)
  double* qpart = nullptr;
  double* qproxy = nullptr;
  double* phipart[3] = {};
  double* phiproxy[3] = {};
  double* disppart[3] = {};
  double* dispproxy[3] = {};
  double* vstage[9] = {};
  double* newvor = nullptr;
  double* dep[3] = {};
  double* vor_next = nullptr;
  double* tr_next = nullptr;
};

struct sbslpm_ctx {
  sbslpm_params prm{};
  Tree tree;
  DevTree dt{};
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  int device = 0;
  std::string err;

  char* arena = nullptr;
  TargetTile* tiles = nullptr;
  int* zero_rows = nullptr;
  int nzero_rows = 0;
  SourceSet* sets = nullptr;
  int* set_start = nullptr;
  int ntiles = 0;
  int nonleaf_tiles = 0;

  // workspace
  SrcRec64* rec64 = nullptr;
  SrcRec32* rec32 = nullptr;
  SrcRecMfma* recmf = nullptr;
  double* qpart = nullptr;
  double* qproxy = nullptr;
  double* phipart[3] = {};
  double* phiproxy[3] = {};
  double* disppart[3] = {};
  double* dispproxy[3] = {};
  double* vstage[9] = {};  // RK4 stages 1..3
  double* newvor = nullptr;
  double* dep[3] = {};
  double* vor_next = nullptr;
  double* tr_next = nullptr;
  double* scratch = nullptr;  // direct sums, grown on demand
  int64_t scratch_len = 0;
  // adjoint workspace and source-major interaction list, allocated on the
  // first adjoint call
  char* adj_arena = nullptr;
  TargetTile* adj_tiles = nullptr;
  SourceSet* adj_targets = nullptr;
  int* adj_zero_rows = nullptr;
  int adj_ntiles = 0;
  int adj_nzero = 0;
  double* adj_ones = nullptr;
  double* adj_psip[3] = {};  // particle copies of the cotangent (upward-pass output)
  double* adj_psix[3] = {};  // its anterpolation to the proxy rows
  void* adj_trec = nullptr;  // target records
  SrcRec64* adj_srec = nullptr;
  double* adj_qbp = nullptr;  // charge cotangent at source rows: particles, proxies
  double* adj_qbx = nullptr;
  double* adj_qb = nullptr;  // dL/dq_part after the downward pass
  double* adj_dbp[3] = {};   // displacement cotangent: particles, proxies
  double* adj_dbx[3] = {};
  double* adj_v[12] = {};    // stage velocities v0..v3 of the step adjoint
  double* adj_vb[12] = {};   // and their cotangents
  Work w1;                     // the single-member workspace (aliases the arena buffers above)
  Work wens;                   // ensemble workspace, allocated on demand
  char* ens_arena = nullptr;
  hipGraphExec_t ens_graph = nullptr;
  const void* ens_key[10] = {};
  // shallow-water workspace, allocated on the first SWE call
  char* swe_arena = nullptr;
  SrcRecSwe* recswe = nullptr;
  SrcRecSwe32* recswe32 = nullptr;
  double* qdpart = nullptr;
  double* qdproxy = nullptr;
  // shallow-water step workspace (stage displacements, velocities, tendencies,
  // displaced state, surface Laplacian, arrival values), allocated on the first step
  char* sws_arena = nullptr;
  hipGraphExec_t swe_graph = nullptr;  // captured shallow-water step, keyed by the caller's pointers
  const void* swe_graph_key[13] = {};
  struct SweStep {
    double *dpart[3], *dprox[3], *eff, *lap, *v[3][3], *tend[3][3], *st[3], *da, *arr[3], *dep[3], *ntr;
    double *qp, *qx, *pp, *px;
  } sws{};
  // shallow-water step adjoint workspace, allocated on the first SWE adjoint call
  char* swa_arena = nullptr;
  struct SweAdj {
    double *vb[12], *tb[9], *rb[3], *lapbar, *qzp, *qzx, *qdp, *qdx, *qep, *qex, *qz, *qd, *qe, *dbp[3], *dbx[3];
    double *hscat, *hh, *stages;
    void* trec;
    SrcRec64* srec;
  } swa{};

  // Kernel timing (sbslpm_ctx_enable_timing).
  bool timing = false;
  bool capturing = false;
  struct Pending {
    int kernel;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  double kernel_ms[9] = {};
  int64_t kernel_launches[9] = {};

  // Launch shape of the Biot-Savart pair kernels (sbslpm_autotune).
  PairShape shape{};
  // Launch shape and chunks of four members per wave of the multi-RHS kernel
  // (sbslpm_autotune_ensemble); mrhs_cb = 0: untuned, `shape` and the default.
  PairShape mrhs_shape{};
  int mrhs_cb = 0;

  // hipGraph of one BVE step, keyed by the caller's pointers.
  hipGraphExec_t graph = nullptr;
  const void* graph_key[10] = {};
};

// Irregular particle tree on the device (the lat-lon path), bound to a context
// whose stream and error word it uses.
struct sbslpm_irreg {
  sbslpm_ctx* ctx = nullptr;
  IrregTree tree;
  DevTree dt{};
  char* arena = nullptr;
  int* perm = nullptr;
  int* leaf_first = nullptr;
  int* leaf_count = nullptr;
  int* point_leaf = nullptr;
  int* leaves = nullptr;
  int nleaves = 0;
  int* nonleaf = nullptr;
  TargetTile* tiles = nullptr;
  int ntiles = 0;
  SourceSet* sets = nullptr;
  int* zero_rows = nullptr;
  int nzero = 0;
  // workspace: sorted point arrays [N], proxy arrays [P][n]
  double* q[2] = {};
  double* qx[2] = {};
  double* part[3] = {};
  double* prox[3] = {};
  SrcRecSwe* recswe = nullptr;
  SrcRec64* rec64 = nullptr;
  double* w = nullptr;  // [N][2 (d+1)] BLI basis of each sorted point on its leaf
  // ancestors of leaf i: anc[3 e .. 3 e + 2] = (panel, xi matrix, eta matrix of
  // anc_tab) for e in [anc_first[i], anc_first[i + 1]): the one-launch L2L sweep
  int* anc_first = nullptr;
  int* anc = nullptr;
  double* anc_tab = nullptr;
  // lat-lon step workspace (tendencies, arrival values, pole means), allocated on
  // the first step
  double* ll_ws = nullptr;
};

namespace {

#define HIP_TRY(ctx, expr)                                                                  \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      if (ctx) (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);              \
      return SBSLPM_E_HIP;                                                                  \
    }                                                                                       \
  } while (0)

int fail(sbslpm_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

// Checks the launch status of the last kernels.
int launched(sbslpm_ctx* c) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, SBSLPM_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return SBSLPM_OK;
}

struct Arena {
  size_t off = 0;
  template <typename T>
  size_t take(size_t count) {
    size_t o = (off + 255) & ~size_t(255);
    off = o + count * sizeof(T);
    return o;
  }
};

int upload(sbslpm_ctx* c, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return SBSLPM_OK;
  HIP_TRY(c, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return SBSLPM_OK;
}

// A launch shape for the context's full target-major tile list: with the
// cost-balanced XCD mapping (xcd 2) the grid is 8 x the largest XCD range in
// workgroups of `waves` tiles, for the ranges of the matrix-core kernels
// (mfma) or of the vector-unit ones.
static PairShape full_list_shape(const Tree& t, int waves, int xcd, bool mfma) {
  PairShape sh{waves, xcd, 0};
  if (xcd == 2) {
    const int* xs = t.xcd_start.data() + (mfma ? 0 : 9);
    int most = 0;
    for (int x = 0; x < 8; ++x) most = std::max(most, (xs[x + 1] - xs[x] + waves - 1) / waves);
    sh.grid = 8 * most;
  }
  return sh;
}

// The logarithm and exponential tables of the FP64 pair kernels, built once
// per process.
struct HostTables {
  std::vector<LogTabEntry> log = std::vector<LogTabEntry>(kLogTabSize);
  std::vector<double> exp = std::vector<double>(kExpTabSize);
  std::vector<double> exp_neg = std::vector<double>(kExpNegTabSize);
  HostTables() {
    build_log_table(log.data());
    build_exp_table(exp.data());
    build_exp_neg_table(exp_neg.data());
  }
};
const HostTables& host_tables() {
  static const HostTables tables;
  return tables;
}

int velocity(sbslpm_ctx* c, const Work& w, int nmem, const double* area, const double* vor,
             const double* const* dpart, const double* const* dprox, double* const* vel);

// Brackets one pair-kernel launch with events when timing is on.
struct KernelTimer {
  sbslpm_ctx* c;
  int kernel;
  hipEvent_t a = nullptr, b = nullptr;
  KernelTimer(sbslpm_ctx* c_, int k) : c(c_), kernel(k) {
    if (c->timing && !c->capturing && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
      hipEventRecord(a, c->stream);
    else
      a = b = nullptr;
  }
  ~KernelTimer() {
    if (!a) return;
    hipEventRecord(b, c->stream);
    c->pending.push_back({kernel, a, b});
  }
};

}  // namespace

extern "C" {

int sbslpm_ctx_create(const sbslpm_params* p, sbslpm_ctx** out) {
  if (!p || !out) return SBSLPM_E_ARG;
  *out = nullptr;
  sbslpm_ctx* c = new (std::nothrow) sbslpm_ctx();
  if (!c) return SBSLPM_E_NOMEM;
  c->prm = *p;
  if (p->precision < SBSLPM_FP64 || p->precision > SBSLPM_MIXED_VALU) {
    delete c;
    return SBSLPM_E_ARG;
  }
  if (p->rk4_combine != SBSLPM_RK4_REFERENCE && p->rk4_combine != SBSLPM_RK4_TEXTBOOK) {
    delete c;
    return SBSLPM_E_ARG;
  }
  if (p->tracer_count < 0 || !(p->kernel_eps >= 0)) {
    delete c;
    return SBSLPM_E_ARG;
  }
  if (p->topography < SBSLPM_TOPO_NONE || p->topography > SBSLPM_TOPO_GAUSSIAN || !(p->gravity >= 0) ||
      (p->swe_form & ~(SBSLPM_SWE_FORM_REG_LAPLACIAN | SBSLPM_SWE_FORM_SURFACE_CORRECTION)) ||
      (p->topography == SBSLPM_TOPO_GAUSSIAN && !(p->mountain_width > 0))) {
    delete c;
    return SBSLPM_E_ARG;
  }
  TreeConfig cfg{p->levels, p->interp_degree, p->leaf_degree, p->theta, p->weight_rule};
  std::string msg = build_tree(cfg, c->tree);
  if (!msg.empty()) {
    delete c;
    return SBSLPM_E_ARG;
  }
  const Tree& t = c->tree;
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return SBSLPM_E_HIP;
  }
  // device address of the fixed scratch of sbslpm_direct_sum_inv_lap_noalloc,
  // looked up once per device while it is current
  if (direct_noalloc_init(c->device) != 0) {
    delete c;
    return SBSLPM_E_HIP;
  }

  // Device copies of the constant tables + workspace in one allocation.
  const size_t Pn = (size_t)t.P * t.n, NP = (size_t)t.npart, U = (size_t)t.U;
  const int ntr = std::max(p->tracer_count, 1);
  std::vector<double> hiT(t.n), loT(t.n), p2mT(t.dp1 * t.m);
  for (int a = 0; a < t.dp1; ++a)
    for (int k = 0; k < t.dp1; ++k) {
      hiT[k * t.dp1 + a] = t.m2m_hi[a * t.dp1 + k];
      loT[k * t.dp1 + a] = t.m2m_lo[a * t.dp1 + k];
    }
  for (int a = 0; a < t.dp1; ++a)
    for (int k = 0; k < t.m; ++k) p2mT[k * t.dp1 + a] = t.p2m[a * t.m + k];

  Arena A;
  size_t o_pan = A.take<DevPanel>(t.P), o_px = A.take<double>(U), o_py = A.take<double>(U), o_pz = A.take<double>(U);
  size_t o_chd = A.take<double>(t.dp1), o_chm = A.take<double>(t.m);
  size_t o_hi = A.take<double>(t.n), o_lo = A.take<double>(t.n), o_hiT = A.take<double>(t.n),
         o_loT = A.take<double>(t.n);
  size_t o_p2m = A.take<double>(t.dp1 * t.m), o_p2mT = A.take<double>(t.dp1 * t.m);
  size_t o_anc = A.take<double>(std::max<size_t>(t.anc_l2l.size(), 1));
  size_t o_l2phi = A.take<double>(t.l2p_hi.size()), o_l2plo = A.take<double>(t.l2p_lo.size());
  size_t o_l2pThi = A.take<double>(t.l2pT_hi.size()), o_l2pTlo = A.take<double>(t.l2pT_lo.size());
  size_t o_pl = A.take<int>(t.proxy_leaf.size()), o_pxe = A.take<double>(t.proxy_xieta.size());
  size_t o_po = A.take<int>(t.proxy_order.size());
  size_t o_err = A.take<int>(4), o_zr = A.take<int>(t.zero_rows.size());
  size_t o_tiles = A.take<TargetTile>(t.tiles.size()), o_sets = A.take<SourceSet>(t.srcsets.size());
  size_t o_ss = A.take<int>(t.set_start.size());
  size_t o_rec64 = A.take<SrcRec64>(U), o_rec32 = A.take<SrcRec32>(U), o_recmf = A.take<SrcRecMfma>(t.U16);
  size_t o_qpart = A.take<double>(NP), o_qproxy = A.take<double>(Pn);
  size_t o_phipart = A.take<double>(3 * NP), o_phiproxy = A.take<double>(3 * Pn);
  size_t o_disppart = A.take<double>(3 * NP), o_dispproxy = A.take<double>(3 * Pn);
  size_t o_vstage = A.take<double>(9 * NP), o_newvor = A.take<double>(NP), o_dep = A.take<double>(3 * NP);
  size_t o_vnext = A.take<double>(NP), o_trnext = A.take<double>(NP * ntr);
  size_t o_ltab = A.take<LogTabEntry>(kLogTabSize), o_etab = A.take<double>(kExpTabSize),
         o_entab = A.take<double>(kExpNegTabSize), o_xcd = A.take<int>(18);
  size_t total = A.take<char>(256);
  if (hipMalloc(&c->arena, total) != hipSuccess) {
    delete c;
    return SBSLPM_E_NOMEM;
  }
  char* b = c->arena;
  int rc = SBSLPM_OK;
  rc |= upload(c, b + o_pan, t.dev.data(), t.P * sizeof(DevPanel));
  rc |= upload(c, b + o_px, t.px.data(), U * sizeof(double));
  rc |= upload(c, b + o_py, t.py.data(), U * sizeof(double));
  rc |= upload(c, b + o_pz, t.pz.data(), U * sizeof(double));
  rc |= upload(c, b + o_chd, t.cheb_d.data(), t.dp1 * sizeof(double));
  rc |= upload(c, b + o_chm, t.cheb_m.data(), t.m * sizeof(double));
  rc |= upload(c, b + o_hi, t.m2m_hi.data(), t.n * sizeof(double));
  rc |= upload(c, b + o_lo, t.m2m_lo.data(), t.n * sizeof(double));
  rc |= upload(c, b + o_hiT, hiT.data(), t.n * sizeof(double));
  rc |= upload(c, b + o_loT, loT.data(), t.n * sizeof(double));
  rc |= upload(c, b + o_p2m, t.p2m.data(), t.p2m.size() * sizeof(double));
  rc |= upload(c, b + o_p2mT, p2mT.data(), p2mT.size() * sizeof(double));
  rc |= upload(c, b + o_anc, t.anc_l2l.data(), t.anc_l2l.size() * sizeof(double));
  rc |= upload(c, b + o_l2phi, t.l2p_hi.data(), t.l2p_hi.size() * sizeof(double));
  rc |= upload(c, b + o_l2plo, t.l2p_lo.data(), t.l2p_lo.size() * sizeof(double));
  rc |= upload(c, b + o_l2pThi, t.l2pT_hi.data(), t.l2pT_hi.size() * sizeof(double));
  rc |= upload(c, b + o_l2pTlo, t.l2pT_lo.data(), t.l2pT_lo.size() * sizeof(double));
  rc |= upload(c, b + o_pl, t.proxy_leaf.data(), t.proxy_leaf.size() * sizeof(int));
  rc |= upload(c, b + o_pxe, t.proxy_xieta.data(), t.proxy_xieta.size() * sizeof(double));
  rc |= upload(c, b + o_po, t.proxy_order.data(), t.proxy_order.size() * sizeof(int));
  rc |= upload(c, b + o_tiles, t.tiles.data(), t.tiles.size() * sizeof(TargetTile));
  rc |= upload(c, b + o_sets, t.srcsets.data(), t.srcsets.size() * sizeof(SourceSet));
  rc |= upload(c, b + o_zr, t.zero_rows.data(), t.zero_rows.size() * sizeof(int));
  rc |= upload(c, b + o_ss, t.set_start.data(), t.set_start.size() * sizeof(int));
  {
    const HostTables& ht = host_tables();
    rc |= upload(c, b + o_ltab, ht.log.data(), kLogTabSize * sizeof(LogTabEntry));
    rc |= upload(c, b + o_etab, ht.exp.data(), kExpTabSize * sizeof(double));
    rc |= upload(c, b + o_entab, ht.exp_neg.data(), kExpNegTabSize * sizeof(double));
  }
  rc |= upload(c, b + o_xcd, t.xcd_start.data(), 18 * sizeof(int));
  if (rc != SBSLPM_OK || hipMemset(b + o_err, 0, 4 * sizeof(int)) != hipSuccess) {
    hipFree(c->arena);
    delete c;
    return SBSLPM_E_HIP;
  }
  DevTree& d = c->dt;
  d.panels = (const DevPanel*)(b + o_pan);
  d.px = (const double*)(b + o_px); d.py = (const double*)(b + o_py); d.pz = (const double*)(b + o_pz);
  d.cheb_d = (const double*)(b + o_chd); d.cheb_m = (const double*)(b + o_chm);
  d.m2m_hi = (const double*)(b + o_hi); d.m2m_lo = (const double*)(b + o_lo);
  d.m2m_hiT = (const double*)(b + o_hiT); d.m2m_loT = (const double*)(b + o_loT);
  d.p2m = (const double*)(b + o_p2m); d.p2mT = (const double*)(b + o_p2mT);
  d.anc_l2l = (const double*)(b + o_anc);
  d.l2p_hi = (const double*)(b + o_l2phi);
  d.l2p_lo = (const double*)(b + o_l2plo);
  d.l2pT_hi = (const double*)(b + o_l2pThi);
  d.l2pT_lo = (const double*)(b + o_l2pTlo);
  d.proxy_leaf = (const int*)(b + o_pl); d.proxy_xieta = (const double*)(b + o_pxe);
  d.proxy_order = (const int*)(b + o_po);
  d.logtab = (const LogTabEntry*)(b + o_ltab);
  d.exptab = (const double*)(b + o_etab);
  d.expnegtab = (const double*)(b + o_entab);
  d.xcd_start = (const int*)(b + o_xcd);
  d.err = (int*)(b + o_err);
  d.bf16_m2l = (c->prm.flags & SBSLPM_FLAG_BF16_M2L) ? 1 : 0;
  d.P = t.P; d.Pl = t.Pl; d.offset = t.offset; d.n = t.n; d.dp1 = t.dp1; d.m = t.m; d.m2 = t.m2;
  d.levels = t.cfg.levels; d.npart = t.npart; d.U = t.U;
  d.m2p = t.m2p; d.np16 = t.np16; d.U16 = t.U16;
  c->tiles = (TargetTile*)(b + o_tiles);
  c->sets = (SourceSet*)(b + o_sets);
  c->ntiles = (int)t.tiles.size();
  c->nonleaf_tiles = t.nonleaf_tiles;
  c->set_start = (int*)(b + o_ss);
  c->zero_rows = (int*)(b + o_zr);
  c->nzero_rows = (int)t.zero_rows.size();
  c->rec64 = (SrcRec64*)(b + o_rec64);
  c->rec32 = (SrcRec32*)(b + o_rec32);
  c->recmf = (SrcRecMfma*)(b + o_recmf);
  c->qpart = (double*)(b + o_qpart);
  c->qproxy = (double*)(b + o_qproxy);
  for (int f = 0; f < 3; ++f) {
    c->phipart[f] = (double*)(b + o_phipart) + f * NP;
    c->phiproxy[f] = (double*)(b + o_phiproxy) + f * Pn;
    c->disppart[f] = (double*)(b + o_disppart) + f * NP;
    c->dispproxy[f] = (double*)(b + o_dispproxy) + f * Pn;
    c->dep[f] = (double*)(b + o_dep) + f * NP;
  }
  for (int k = 0; k < 9; ++k) c->vstage[k] = (double*)(b + o_vstage) + k * NP;
  c->newvor = (double*)(b + o_newvor);
  c->vor_next = (double*)(b + o_vnext);
  c->tr_next = (double*)(b + o_trnext);
  {
    Work& w = c->w1;
    w.members = 1;
    w.rec64 = c->rec64; w.rec32 = c->rec32; w.recmf = c->recmf;
    w.qpart = c->qpart; w.qproxy = c->qproxy;
    for (int f = 0; f < 3; ++f) {
      w.phipart[f] = c->phipart[f]; w.phiproxy[f] = c->phiproxy[f];
      w.disppart[f] = c->disppart[f]; w.dispproxy[f] = c->dispproxy[f]; w.dep[f] = c->dep[f];
    }
    for (int k = 0; k < 9; ++k) w.vstage[k] = c->vstage[k];
    w.newvor = c->newvor; w.vor_next = c->vor_next; w.tr_next = c->tr_next;
  }
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    hipFree(c->arena);
    delete c;
    return SBSLPM_E_HIP;
  }
  c->stream = c->own_stream;
  *out = c;
  return SBSLPM_OK;
}

void sbslpm_ctx_destroy(sbslpm_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->graph) hipGraphExecDestroy(c->graph);
  for (auto& p : c->pending) {
    hipEventDestroy(p.a);
    hipEventDestroy(p.b);
  }
  if (c->scratch) hipFree(c->scratch);
  if (c->swe_arena) hipFree(c->swe_arena);
  if (c->swe_graph) hipGraphExecDestroy(c->swe_graph);
  if (c->sws_arena) hipFree(c->sws_arena);
  if (c->swa_arena) hipFree(c->swa_arena);
  if (c->adj_arena) hipFree(c->adj_arena);
  if (c->ens_graph) hipGraphExecDestroy(c->ens_graph);
  if (c->ens_arena) hipFree(c->ens_arena);
  if (c->arena) hipFree(c->arena);
  if (c->own_stream) hipStreamDestroy(c->own_stream);
  delete c;
}

int sbslpm_ctx_set_stream(sbslpm_ctx* c, void* stream) {
  if (!c) return SBSLPM_E_ARG;
  c->stream = (hipStream_t)stream;  // NULL: the default (legacy) stream
  if (c->graph) {
    hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  return SBSLPM_OK;
}

int sbslpm_ctx_set_precision(sbslpm_ctx* c, int precision) {
  if (!c || precision < SBSLPM_FP64 || precision > SBSLPM_MIXED_VALU) return SBSLPM_E_ARG;
  if (c->capturing) return fail(c, SBSLPM_E_ARG, "set_precision: not allowed during graph capture");
  if (precision != c->prm.precision) {
    // captured steps hold the old precision's kernels; the ensemble workspace
    // holds its record kind only (ensure_ensemble reallocates it on next use)
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    hipGraphExec_t* graphs[3] = {&c->graph, &c->ens_graph, &c->swe_graph};
    for (hipGraphExec_t* g : graphs)
      if (*g) {
        hipGraphExecDestroy(*g);
        *g = nullptr;
      }
  }
  c->prm.precision = precision;
  // the balanced XCD ranges (and so the grid) differ between matrix-core and vector-unit kernels
  c->shape = full_list_shape(c->tree, c->shape.waves, c->shape.xcd, precision == SBSLPM_MIXED);
  return SBSLPM_OK;
}

const char* sbslpm_last_error(const sbslpm_ctx* c) { return c ? c->err.c_str() : "null context"; }

int sbslpm_synchronize(sbslpm_ctx* c) {
  if (!c) return SBSLPM_E_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  int flag = 0;
  HIP_TRY(c, hipMemcpy(&flag, c->dt.err, sizeof(int), hipMemcpyDeviceToHost));
  if (flag) {
    HIP_TRY(c, hipMemset(c->dt.err, 0, sizeof(int)));
    return fail(c, SBSLPM_E_GEOMETRY,
                (flag & kErrLocate) ? "point location failed (non-finite or off-sphere point)"
                                    : "interpolation point outside its panel");
  }
  return SBSLPM_OK;
}

int sbslpm_ctx_enable_timing(sbslpm_ctx* c, int enable) {
  if (!c) return SBSLPM_E_ARG;
  c->timing = enable != 0;
  return SBSLPM_OK;
}

int sbslpm_kernel_time(sbslpm_ctx* c, int kernel, double* ms, int64_t* launches, int reset) {
  if (!c || kernel < 0 || kernel > 8) return SBSLPM_E_ARG;
  for (auto& p : c->pending) {
    HIP_TRY(c, hipEventSynchronize(p.b));
    float t = 0;
    HIP_TRY(c, hipEventElapsedTime(&t, p.a, p.b));
    c->kernel_ms[p.kernel] += t;
    c->kernel_launches[p.kernel] += 1;
    hipEventDestroy(p.a);
    hipEventDestroy(p.b);
  }
  c->pending.clear();
  if (ms) *ms = c->kernel_ms[kernel];
  if (launches) *launches = c->kernel_launches[kernel];
  if (reset) {
    c->kernel_ms[kernel] = 0;
    c->kernel_launches[kernel] = 0;
  }
  return SBSLPM_OK;
}

int sbslpm_get_sizes(const sbslpm_ctx* c, int64_t* s) {
  if (!c || !s) return SBSLPM_E_ARG;
  const Tree& t = c->tree;
  s[0] = t.P; s[1] = t.Pl; s[2] = t.n; s[3] = t.m2; s[4] = t.interaction_count(); s[5] = t.npart;
  s[6] = t.offset; s[7] = (int64_t)t.tiles.size();
  return SBSLPM_OK;
}

int sbslpm_spec_points(const sbslpm_ctx* c, double* x, double* y, double* z, double* area) {
  if (!c) return SBSLPM_E_ARG;
  const Tree& t = c->tree;
  size_t bytes = (size_t)t.npart * sizeof(double);
  if (x) std::memcpy(x, t.spec_x.data(), bytes);
  if (y) std::memcpy(y, t.spec_y.data(), bytes);
  if (z) std::memcpy(z, t.spec_z.data(), bytes);
  if (area) std::memcpy(area, t.spec_area.data(), bytes);
  return SBSLPM_OK;
}

int sbslpm_interaction_list(const sbslpm_ctx* c, int* list) {
  if (!c || !list) return SBSLPM_E_ARG;
  std::memcpy(list, c->tree.ilist.data(), c->tree.ilist.size() * sizeof(int));
  return SBSLPM_OK;
}

int sbslpm_upward_pass(sbslpm_ctx* c, const double* area, const double* pots, double* q_part, double* q_proxy) {
  if (!c || !area || !pots || !q_part || !q_proxy) return fail(c, SBSLPM_E_ARG, "upward_pass: null pointer");
  launch_upward_pass(c->dt, area, pots, q_part, q_proxy, c->stream);
  return launched(c);
}

int sbslpm_bve_vel_interactions(sbslpm_ctx* c, const double* dpx, const double* dpy, const double* dpz,
                                const double* dxx, const double* dxy, const double* dxz, const double* q_part,
                                const double* q_proxy, double* ppx, double* ppy, double* ppz, double* pxx,
                                double* pxy, double* pxz) {
  if (!c || !q_part || !q_proxy || !ppx || !ppy || !ppz || !pxx || !pxy || !pxz)
    return fail(c, SBSLPM_E_ARG, "bve_vel_interactions: null pointer");
  const double* dp[3] = {dpx, dpy, dpz};
  const double* dx[3] = {dxx, dxy, dxz};
  double* op[3] = {ppx, ppy, ppz};
  double* ox[3] = {pxx, pxy, pxz};
  const double* const* dpp = dpx ? dp : nullptr;
  const double* const* dxp = dxx ? dx : nullptr;
  if (c->prm.precision == SBSLPM_MIXED) {
    launch_source_records_mfma(c->dt, q_part, q_proxy, dpp, dxp, c->recmf, c->stream);
    KernelTimer tm(c, SBSLPM_KERNEL_BVE_MIXED);
    launch_bve_interact_mfma(c->dt, c->tiles, c->ntiles, c->sets, c->recmf, c->prm.kernel_eps, op, ox, c->stream);
  } else if (c->prm.precision == SBSLPM_MIXED_VALU) {
    launch_source_records32(c->dt, q_part, q_proxy, dpp, dxp, c->rec32, c->stream);
    KernelTimer tm(c, SBSLPM_KERNEL_BVE_MIXED);
    launch_bve_interact_mixed(c->dt, c->tiles, c->ntiles, c->sets, c->rec32, c->prm.kernel_eps, op, ox, c->stream);
  } else {
    launch_source_records64(c->dt, q_part, q_proxy, dpp, dxp, true, c->rec64, c->stream);
    KernelTimer tm(c, SBSLPM_KERNEL_BVE_FP64);
    launch_bve_interact64(c->dt, c->tiles, c->ntiles, c->sets, c->rec64, c->prm.kernel_eps, op, ox, c->stream);
  }
  return launched(c);
}

// Source records and the log pair sums over the full list, accumulated into
// the target rows or (assign) written over them.
static void poisson_pairs(sbslpm_ctx* c, const double* q_part, const double* q_proxy, double* psi_part,
                          double* psi_proxy, bool assign) {
  if (c->prm.precision != SBSLPM_FP64) {
    launch_source_records32(c->dt, q_part, q_proxy, nullptr, nullptr, c->rec32, c->stream);
    KernelTimer tm(c, SBSLPM_KERNEL_LOG_MIXED);
    launch_poisson_interact_mixed(c->dt, c->tiles, c->ntiles, c->sets, c->rec32, psi_part, psi_proxy, c->stream,
                                  assign);
  } else {
    // the Poisson kernel reads the node coordinates unprojected (inverse_poisson_impl.hpp:42-46)
    launch_source_records64(c->dt, q_part, q_proxy, nullptr, nullptr, false, c->rec64, c->stream);
    KernelTimer tm(c, SBSLPM_KERNEL_LOG_FP64);
    launch_poisson_interact64(c->dt, c->tiles, c->ntiles, c->sets, c->rec64, psi_part, psi_proxy, c->stream,
                              assign);
  }
}

int sbslpm_poisson_fmm_interactions(sbslpm_ctx* c, const double* q_part, const double* q_proxy, double* psi_part,
                                    double* psi_proxy) {
  if (!c || !q_part || !q_proxy || !psi_part || !psi_proxy)
    return fail(c, SBSLPM_E_ARG, "poisson_fmm_interactions: null pointer");
  poisson_pairs(c, q_part, q_proxy, psi_part, psi_proxy, false);
  return launched(c);
}

int sbslpm_downward_pass(sbslpm_ctx* c, double* phi_proxy, const double* phi_part, double* sols) {
  if (!c || !phi_proxy || !sols) return fail(c, SBSLPM_E_ARG, "downward_pass: null pointer");
  double* px[1] = {phi_proxy};
  const double* pp[1] = {phi_part};
  double* so[1] = {sols};
  launch_downward_pass(c->dt, 1, px, pp, so, c->stream);
  return launched(c);
}

int sbslpm_downward_pass_3(sbslpm_ctx* c, double* p1, double* p2, double* p3, const double* q1, const double* q2,
                           const double* q3, double* s1, double* s2, double* s3) {
  if (!c || !p1 || !p2 || !p3 || !s1 || !s2 || !s3) return fail(c, SBSLPM_E_ARG, "downward_pass_3: null pointer");
  double* px[3] = {p1, p2, p3};
  const double* pp[3] = {q1, q2, q3};
  double* so[3] = {s1, s2, s3};
  launch_downward_pass(c->dt, 3, px, pp, so, c->stream);
  return launched(c);
}

int sbslpm_bve_velocity(sbslpm_ctx* c, const double* area, const double* vor, const double* dpx, const double* dpy,
                        const double* dpz, const double* dxx, const double* dxy, const double* dxz, double* vx,
                        double* vy, double* vz) {
  if (!c || !area || !vor || !vx || !vy || !vz) return fail(c, SBSLPM_E_ARG, "bve_velocity: null pointer");
  const double* dp[3] = {dpx, dpy, dpz};
  const double* dx[3] = {dxx, dxy, dxz};
  double* v[3] = {vx, vy, vz};
  return velocity(c, c->w1, 1, area, vor, dpx ? dp : nullptr, dxx ? dx : nullptr, v);
}

int sbslpm_poisson_solve(sbslpm_ctx* c, const double* area, const double* f, double* psi) {
  if (!c || !area || !f || !psi) return fail(c, SBSLPM_E_ARG, "poisson_solve: null pointer");
  // The reference's sequence (upward pass, interactions into zeroed rows,
  // downward pass) with the fused passes: the leaf proxy rows are neither
  // interaction sources nor targets, so the upward pass leaves them unwritten
  // and the downward pass goes from the non-leaf rows to the particles; the
  // pair sums assign their target rows and the non-leaf rows without tiles
  // are zeroed.
  launch_upward_fused(c->dt, area, f, c->qpart, c->qproxy, c->stream);
  launch_zero_rows(c->zero_rows, c->nzero_rows, c->tree.n, 1, c->phiproxy, c->stream);
  poisson_pairs(c, c->qpart, c->qproxy, c->phipart[0], c->phiproxy[0], true);
  double* px[1] = {c->phiproxy[0]};
  const double* pp[1] = {c->phipart[0]};
  double* so[1] = {psi};
  launch_downward_pass(c->dt, 1, px, pp, so, c->stream, true);
  return launched(c);
}

// Shallow-water workspace: records [U] and the divergence charges.
static int ensure_swe(sbslpm_ctx* c) {
  if (c->swe_arena) return SBSLPM_OK;
  const Tree& t = c->tree;
  Arena A;
  const size_t o_rec = A.take<SrcRecSwe>(t.U), o_rec32 = A.take<SrcRecSwe32>(t.U), o_qp = A.take<double>(t.npart),
               o_qx = A.take<double>((size_t)t.P * t.n), total = A.take<char>(256);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (hipMalloc(&c->swe_arena, total) != hipSuccess) {
    c->swe_arena = nullptr;
    return fail(c, SBSLPM_E_NOMEM, "shallow-water workspace: out of device memory");
  }
  c->recswe = (SrcRecSwe*)(c->swe_arena + o_rec);
  c->recswe32 = (SrcRecSwe32*)(c->swe_arena + o_rec32);
  c->qdpart = (double*)(c->swe_arena + o_qp);
  c->qdproxy = (double*)(c->swe_arena + o_qx);
  return SBSLPM_OK;
}

static int swe_interactions(sbslpm_ctx* c, int variant, const double* dpx, const double* dpy, const double* dpz,
                            const double* dxx, const double* dxy, const double* dxz, const double* qz_part,
                            const double* qz_proxy, const double* qd_part, const double* qd_proxy, double* const* op,
                            double* const* ox, bool assign) {
  int rc = ensure_swe(c);
  if (rc) return rc;
  const double* dp[3] = {dpx, dpy, dpz};
  const double* dx[3] = {dxx, dxy, dxz};
  if (c->prm.precision != SBSLPM_FP64 && variant != 2) {
    // FP32 pairs on the vector unit for variants 1 and 3 (variant 2 stays FP64)
    launch_source_records_swe32(c->dt, variant, qz_part, qz_proxy, qd_part, qd_proxy, dpx ? dp : nullptr,
                                dxx ? dx : nullptr, c->recswe32, c->stream);
    KernelTimer tm(c, SBSLPM_KERNEL_SWE_MIXED);
    launch_swe_interact_mixed(c->dt, c->tiles, c->ntiles, c->sets, c->recswe32, c->prm.kernel_eps, op, ox, c->stream,
                              assign);
    return launched(c);
  }
  launch_source_records_swe(c->dt, variant, qz_part, qz_proxy, qd_part, qd_proxy, dpx ? dp : nullptr,
                            dxx ? dx : nullptr, c->recswe, c->stream);
  {
    KernelTimer tm(c, SBSLPM_KERNEL_SWE_FP64);
    launch_swe_interact64(c->dt, variant, c->tiles, c->ntiles, c->sets, c->recswe, c->prm.kernel_eps, op, ox,
                          c->stream, assign);
  }
  return launched(c);
}

int sbslpm_swe_vel_interactions(sbslpm_ctx* c, const double* dpx, const double* dpy, const double* dpz,
                                const double* dxx, const double* dxy, const double* dxz, const double* qz_part,
                                const double* qz_proxy, const double* qd_part, const double* qd_proxy, double* ppx,
                                double* ppy, double* ppz, double* pxx, double* pxy, double* pxz) {
  if (!c || !qz_part || !qz_proxy || !qd_part || !qd_proxy || !ppx || !ppy || !ppz || !pxx || !pxy || !pxz)
    return fail(c, SBSLPM_E_ARG, "swe_vel_interactions: null pointer");
  double* op[3] = {ppx, ppy, ppz};
  double* ox[3] = {pxx, pxy, pxz};
  return swe_interactions(c, 1, dpx, dpy, dpz, dxx, dxy, dxz, qz_part, qz_proxy, qd_part, qd_proxy, op, ox, false);
}

int sbslpm_swe_vel_interactions_2(sbslpm_ctx* c, const double* dpx, const double* dpy, const double* dpz,
                                  const double* dxx, const double* dxy, const double* dxz, const double* qz_part,
                                  const double* qz_proxy, const double* qd_part, const double* qd_proxy, double* ppx,
                                  double* ppy, double* ppz, double* pxx, double* pxy, double* pxz) {
  if (!c || !qz_part || !qz_proxy || !qd_part || !qd_proxy || !ppx || !ppy || !ppz || !pxx || !pxy || !pxz)
    return fail(c, SBSLPM_E_ARG, "swe_vel_interactions_2: null pointer");
  double* op[3] = {ppx, ppy, ppz};
  double* ox[3] = {pxx, pxy, pxz};
  return swe_interactions(c, 2, dpx, dpy, dpz, dxx, dxy, dxz, qz_part, qz_proxy, qd_part, qd_proxy, op, ox, false);
}

int sbslpm_swe_velocity(sbslpm_ctx* c, int variant, const double* area, const double* vor, const double* div,
                        const double* dpx, const double* dpy, const double* dpz, const double* dxx, const double* dxy,
                        const double* dxz, double* vx, double* vy, double* vz) {
  if (!c || !area || !vor || !div || !vx || !vy || !vz || (variant != 1 && variant != 2))
    return fail(c, SBSLPM_E_ARG, "swe_velocity: bad argument");
  int rc = ensure_swe(c);
  if (rc) return rc;
  const UpCharge vd[2] = {{area, vor, c->qpart, c->qproxy}, {area, div, c->qdpart, c->qdproxy}};
  launch_upward_sets(c->dt, vd, 2, c->stream);
  launch_zero_rows(c->zero_rows, c->nzero_rows, c->tree.n, 3, c->phiproxy, c->stream);
  rc = swe_interactions(c, variant, dpx, dpy, dpz, dxx, dxy, dxz, c->qpart, c->qproxy, c->qdpart, c->qdproxy,
                        c->phipart, c->phiproxy, true);
  if (rc) return rc;
  const double* pp[3] = {c->phipart[0], c->phipart[1], c->phipart[2]};
  double* vel[3] = {vx, vy, vz};
  launch_downward_pass(c->dt, 3, c->phiproxy, pp, vel, c->stream, true);
  return launched(c);
}

// Poisson (kind 0) or Biot-Savart (kind 1) FMM at arbitrary targets, FP64:
// upward pass, pair sums for the non-leaf target panels only, L2L sweep
// (the last level assigns the leaf rows), then one wave per target.
static int eval_targets(sbslpm_ctx* c, int kind, const double* area, const double* f, int64_t nt, const double* tx,
                        const double* ty, const double* tz, double* const* out) {
  const int nc = kind == 0 ? 1 : 3;
  launch_upward_pass(c->dt, area, f, c->qpart, c->qproxy, c->stream);
  launch_zero_rows(c->zero_rows, c->nzero_rows, c->tree.n, nc, c->phiproxy, c->stream);
  // the Poisson kernel reads unprojected nodes, Biot-Savart re-projects (inverse_poisson_impl.hpp:42-46,
  // bve_vel_impl.hpp:54-60)
  launch_source_records64(c->dt, c->qpart, c->qproxy, nullptr, nullptr, kind == 1, c->rec64, c->stream);
  if (kind == 0)
    launch_poisson_interact64(c->dt, c->tiles, c->nonleaf_tiles, c->sets, c->rec64, c->phipart[0], c->phiproxy[0],
                              c->stream, true);
  else
    launch_bve_interact64(c->dt, c->tiles, c->nonleaf_tiles, c->sets, c->rec64, c->prm.kernel_eps, c->phipart,
                          c->phiproxy, c->stream, true);
  launch_l2l(c->dt, nc, c->phiproxy, c->stream, true);
  const double* far[3] = {c->phiproxy[0], c->phiproxy[1], c->phiproxy[2]};
  launch_eval_targets(c->dt, kind, c->set_start, c->sets, c->rec64, far, c->prm.kernel_eps, nt, tx, ty, tz, out,
                      c->stream);
  return launched(c);
}

int sbslpm_poisson_solve_targets(sbslpm_ctx* c, const double* area, const double* f, int64_t nt, const double* tx,
                                 const double* ty, const double* tz, double* psi) {
  if (!c || !area || !f || nt < 0 || (nt > 0 && (!tx || !ty || !tz || !psi)))
    return fail(c, SBSLPM_E_ARG, "poisson_solve_targets: bad argument");
  double* out[3] = {psi, nullptr, nullptr};
  return eval_targets(c, 0, area, f, nt, tx, ty, tz, out);
}

int sbslpm_bve_velocity_targets(sbslpm_ctx* c, const double* area, const double* vor, int64_t nt, const double* tx,
                                const double* ty, const double* tz, double* vx, double* vy, double* vz) {
  if (!c || !area || !vor || nt < 0 || (nt > 0 && (!tx || !ty || !tz || !vx || !vy || !vz)))
    return fail(c, SBSLPM_E_ARG, "bve_velocity_targets: bad argument");
  double* out[3] = {vx, vy, vz};
  return eval_targets(c, 1, area, vor, nt, tx, ty, tz, out);
}

}  // extern "C"

namespace {

int ensure_scratch(sbslpm_ctx* c, int64_t need) {
  if (need <= c->scratch_len) return SBSLPM_OK;
  // growing the buffer synchronizes, allocates and frees the old one, none of
  // which may happen while the stream is captured (the library's own capture
  // or a caller's, e.g. a torch CUDA graph): an already captured graph may
  // still reference the old buffer
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (c->capturing || (c->stream && hipStreamIsCapturing(c->stream, &cs) == hipSuccess &&
                       cs != hipStreamCaptureStatusNone))
    return fail(c, SBSLPM_E_ARG, "scratch buffer would grow during stream capture; call once outside the capture first");
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (c->scratch) hipFree(c->scratch);
  c->scratch = nullptr;
  c->scratch_len = 0;
  HIP_TRY(c, hipMalloc(&c->scratch, need * sizeof(double)));
  c->scratch_len = need;
  return SBSLPM_OK;
}

int velocity(sbslpm_ctx* c, const Work& w, int nmem, const double* area, const double* vor,
             const double* const* dpart, const double* const* dprox, double* const* vel) {
  const Tree& t = c->tree;
  const int64_t Pn = (int64_t)t.P * t.n;
  // Every leaf particle and every proxy of a target panel is written (assigned)
  // by exactly one pair-kernel lane; leaf proxy rows are neither sources nor
  // targets (the fused passes keep them in LDS); only the remaining proxy rows
  // need clearing.
  launch_upward_fused(c->dt, area, vor, w.qpart, w.qproxy, c->stream, nmem);
  launch_zero_rows(c->zero_rows, c->nzero_rows, t.n, 3, w.phiproxy, c->stream, nmem, Pn);
  if (c->prm.precision == SBSLPM_MIXED && nmem > 1 && w.qcol && !dpart && !dprox) {
    // all members at the undisplaced positions: member 0's records give the
    // geometry, the members' charges become columns of one product
    launch_source_records_mfma(c->dt, w.qpart, w.qproxy, nullptr, nullptr, w.recmf, c->stream, 1);
    const int cb = c->mrhs_cb ? c->mrhs_cb : mrhs_chunks_per_wave(nmem);
    launch_source_columns_mrhs(c->dt, w.qpart, w.qproxy, w.recmf, nmem, cb, w.qcol, c->stream);
    KernelTimer tm(c, SBSLPM_KERNEL_BVE_MRHS);
    launch_bve_interact_mrhs(c->dt, c->tiles, c->ntiles, c->sets, w.recmf, w.qcol, nmem, cb, c->prm.kernel_eps,
                             w.phipart, w.phiproxy, c->stream, true, c->mrhs_cb ? c->mrhs_shape : c->shape);
  } else if (c->prm.precision == SBSLPM_MIXED) {
    launch_source_records_mfma(c->dt, w.qpart, w.qproxy, dpart, dprox, w.recmf, c->stream, nmem);
    KernelTimer tm(c, SBSLPM_KERNEL_BVE_MIXED);
    launch_bve_interact_mfma(c->dt, c->tiles, c->ntiles, c->sets, w.recmf, c->prm.kernel_eps, w.phipart,
                             w.phiproxy, c->stream, true, nmem, c->shape);
  } else if (c->prm.precision == SBSLPM_MIXED_VALU) {
    launch_source_records32(c->dt, w.qpart, w.qproxy, dpart, dprox, w.rec32, c->stream, nmem);
    KernelTimer tm(c, SBSLPM_KERNEL_BVE_MIXED);
    launch_bve_interact_mixed(c->dt, c->tiles, c->ntiles, c->sets, w.rec32, c->prm.kernel_eps, w.phipart,
                              w.phiproxy, c->stream, true, nmem, c->shape);
  } else {
    launch_source_records64(c->dt, w.qpart, w.qproxy, dpart, dprox, true, w.rec64, c->stream, nmem);
    KernelTimer tm(c, SBSLPM_KERNEL_BVE_FP64);
    launch_bve_interact64(c->dt, c->tiles, c->ntiles, c->sets, w.rec64, c->prm.kernel_eps, w.phipart, w.phiproxy,
                          c->stream, true, nmem, c->shape);
  }
  const double* pp[3] = {w.phipart[0], w.phipart[1], w.phipart[2]};
  launch_downward_pass(c->dt, 3, w.phiproxy, pp, vel, c->stream, true, nmem);
  return launched(c);
}

// bve_back_rk4_step body (bve_time_step.cpp:318-418) on the context stream,
// for nmem stacked members (tracers only for a single member).
int bve_step_body(sbslpm_ctx* c, const Work& w, int nmem, const double* x, const double* y, const double* z,
                  const double* area, double* vor, double* const* vel0, double* tracers) {
  const Tree& t = c->tree;
  const double dt = c->prm.delta_t, om = c->prm.omega;
  double* const* v1 = w.vstage;
  double* const* v2 = w.vstage + 3;
  double* const* v3 = w.vstage + 6;
  const double* cdp[3] = {w.disppart[0], w.disppart[1], w.disppart[2]};
  const double* cdx[3] = {w.dispproxy[0], w.dispproxy[1], w.dispproxy[2]};
  int rc = velocity(c, w, nmem, area, vor, nullptr, nullptr, vel0);
  const double* stage_in[3][3] = {{vel0[0], vel0[1], vel0[2]}, {v1[0], v1[1], v1[2]}, {v2[0], v2[1], v2[2]}};
  double* const* stage_out[3] = {v1, v2, v3};
  const double taus[3] = {0.5 * dt, 0.5 * dt, dt};
  for (int s = 0; s < 3 && rc == SBSLPM_OK; ++s) {
    launch_bve_displacement(c->dt, x, y, z, stage_in[s], vor, taus[s], om, w.newvor, w.disppart, w.dispproxy,
                            c->stream, nmem);
    rc = velocity(c, w, nmem, area, w.newvor, cdp, cdx, stage_out[s]);
  }
  if (rc) return rc;
  const double* vels[12] = {vel0[0], vel0[1], vel0[2], v1[0], v1[1], v1[2], v2[0], v2[1], v2[2], v3[0], v3[1], v3[2]};
  launch_bve_departure(c->dt, x, y, z, vels, dt, c->prm.rk4_combine, w.dep[0], w.dep[1], w.dep[2], c->stream, nmem);
  const int ntr = nmem == 1 ? c->prm.tracer_count : 0;
  launch_bve_departure_to_target(c->dt, z, w.dep[0], w.dep[1], w.dep[2], vor, tracers, ntr, om, w.vor_next,
                                 w.tr_next, c->stream, nmem);
  launch_copy(vor, w.vor_next, t.npart * nmem, c->stream);
  if (ntr > 0) launch_copy(tracers, w.tr_next, t.npart * ntr, c->stream);
  return launched(c);
}

}  // namespace

extern "C" {

int sbslpm_direct_sum_inv_lap(sbslpm_ctx* c, int64_t n, const double* x, const double* y, const double* z,
                              const double* area, const double* pots, double* soln, int64_t lb, int64_t ub) {
  if (!c || !x || !y || !z || !area || !pots || !soln || n < 0 || lb < 0 || ub > n || lb > ub)
    return fail(c, SBSLPM_E_ARG, "direct_sum_inv_lap: bad argument");
  int64_t nt = ub - lb;
  if (nt == 0) return SBSLPM_OK;
  int rc = ensure_scratch(c, direct_partials_size(nt, n, 1));
  if (rc) return rc;
  launch_direct_inv_lap(nt, x + lb, y + lb, z + lb, n, x, y, z, area, pots, lb, false, c->scratch, soln + lb,
                        c->stream);
  return launched(c);
}

int sbslpm_direct_sum_inv_lap_noalloc(sbslpm_ctx* c, int64_t n, const double* x, const double* y,
                                      const double* z, const double* area, const double* pots, double* soln,
                                      int64_t lb, int64_t ub) {
  if (!c || !x || !y || !z || !area || !pots || !soln || n < 0 || lb < 0 || ub > n || lb > ub)
    return fail(c, SBSLPM_E_ARG, "direct_sum_inv_lap_noalloc: bad argument");
  const int64_t nt = ub - lb;
  if (nt == 0) return SBSLPM_OK;
  // same chunking as sbslpm_direct_sum_inv_lap, so the sums agree to the bit
  const int64_t need = direct_partials_size(nt, n, 1);
  if (need > kNoAllocScratchDoubles)
    return fail(c, SBSLPM_E_ARG,
                "direct_sum_inv_lap_noalloc: " + std::to_string(nt) + " targets need " + std::to_string(need) +
                    " scratch doubles, above the fixed " + std::to_string(kNoAllocScratchDoubles) +
                    "; split [lb, ub) into smaller slices");
  double* buf = direct_noalloc_scratch(c->device);
  if (!buf) return fail(c, SBSLPM_E_HIP, "direct_sum_inv_lap_noalloc: fixed scratch not set up on this device");
  launch_direct_inv_lap(nt, x + lb, y + lb, z + lb, n, x, y, z, area, pots, lb, false, buf, soln + lb, c->stream);
  return launched(c);
}

int sbslpm_direct_sum_inv_lap_targets(sbslpm_ctx* c, int64_t nt, const double* tx, const double* ty,
                                      const double* tz, int64_t ns, const double* sx, const double* sy,
                                      const double* sz, const double* area, const double* pots, double* soln) {
  if (!c || !tx || !ty || !tz || !sx || !sy || !sz || !area || !pots || !soln || nt < 0 || ns < 0)
    return fail(c, SBSLPM_E_ARG, "direct_sum_inv_lap_targets: bad argument");
  if (nt == 0) return SBSLPM_OK;
  int rc = ensure_scratch(c, direct_partials_size(nt, ns, 1));
  if (rc) return rc;
  launch_direct_inv_lap(nt, tx, ty, tz, ns, sx, sy, sz, area, pots, 0, true, c->scratch, soln, c->stream);
  return launched(c);
}

int sbslpm_direct_sum_bve_vel(sbslpm_ctx* c, int64_t n, const double* x, const double* y, const double* z,
                              const double* area, const double* vor, const double* dx, const double* dy,
                              const double* dz, double* vx, double* vy, double* vz, int64_t lb, int64_t ub) {
  if (!c || !x || !y || !z || !area || !vor || !vx || !vy || !vz || n < 0 || lb < 0 || ub > n || lb > ub)
    return fail(c, SBSLPM_E_ARG, "direct_sum_bve_vel: bad argument");
  if ((dx || dy || dz) && !(dx && dy && dz)) return fail(c, SBSLPM_E_ARG, "direct_sum_bve_vel: partial disp");
  int64_t nt = ub - lb;
  if (nt == 0) return SBSLPM_OK;
  int rc = ensure_scratch(c, direct_partials_size(nt, n, 3));
  if (rc) return rc;
  const double* disp[3] = {dx, dy, dz};
  double* vel[3] = {vx + lb, vy + lb, vz + lb};
  launch_direct_bve_vel(nt, x + lb, y + lb, z + lb, n, x, y, z, area, vor, dx ? disp : nullptr, c->prm.kernel_eps,
                        c->scratch, vel, c->stream);
  return launched(c);
}

int sbslpm_direct_sum_bve_vel_targets(sbslpm_ctx* c, int64_t nt, const double* tx, const double* ty,
                                      const double* tz, int64_t ns, const double* sx, const double* sy,
                                      const double* sz, const double* area, const double* vor, double* vx,
                                      double* vy, double* vz) {
  if (!c || !tx || !ty || !tz || !sx || !sy || !sz || !area || !vor || !vx || !vy || !vz || nt < 0 || ns < 0)
    return fail(c, SBSLPM_E_ARG, "direct_sum_bve_vel_targets: bad argument");
  if (nt == 0) return SBSLPM_OK;
  int rc = ensure_scratch(c, direct_partials_size(nt, ns, 3));
  if (rc) return rc;
  double* vel[3] = {vx, vy, vz};
  launch_direct_bve_vel(nt, tx, ty, tz, ns, sx, sy, sz, area, vor, nullptr, c->prm.kernel_eps, c->scratch, vel,
                        c->stream);
  return launched(c);
}

int sbslpm_bve_displacement(sbslpm_ctx* c, const double* x, const double* y, const double* z, const double* vx,
                            const double* vy, const double* vz, const double* vor, double tau, double* new_vor,
                            double* dpx, double* dpy, double* dpz, double* dxx, double* dxy, double* dxz) {
  if (!c || !x || !y || !z || !vx || !vy || !vz || !vor || !new_vor || !dpx || !dpy || !dpz || !dxx || !dxy || !dxz)
    return fail(c, SBSLPM_E_ARG, "bve_displacement: null pointer");
  const double* v[3] = {vx, vy, vz};
  double* dp[3] = {dpx, dpy, dpz};
  double* dxp[3] = {dxx, dxy, dxz};
  launch_bve_displacement(c->dt, x, y, z, v, vor, tau, c->prm.omega, new_vor, dp, dxp, c->stream);
  return launched(c);
}

int sbslpm_bve_find_departure_points(sbslpm_ctx* c, const double* x, const double* y, const double* z,
                                     const double* const* vels, double dt, double* dep_x, double* dep_y,
                                     double* dep_z) {
  if (!c || !x || !y || !z || !vels || !dep_x || !dep_y || !dep_z)
    return fail(c, SBSLPM_E_ARG, "bve_find_departure_points: null pointer");
  for (int i = 0; i < 12; ++i)
    if (!vels[i]) return fail(c, SBSLPM_E_ARG, "bve_find_departure_points: null velocity");
  launch_bve_departure(c->dt, x, y, z, vels, dt, c->prm.rk4_combine, dep_x, dep_y, dep_z, c->stream);
  return launched(c);
}

int sbslpm_bve_departure_to_target(sbslpm_ctx* c, const double* z, const double* dep_x, const double* dep_y,
                                   const double* dep_z, const double* vor, const double* tracers, int ntr,
                                   double* new_vor, double* new_tracers) {
  if (!c || !z || !dep_x || !dep_y || !dep_z || !vor || !new_vor || ntr < 0 || (ntr > 0 && (!tracers || !new_tracers)))
    return fail(c, SBSLPM_E_ARG, "bve_departure_to_target: bad argument");
  launch_bve_departure_to_target(c->dt, z, dep_x, dep_y, dep_z, vor, tracers, ntr, c->prm.omega, new_vor,
                                 new_tracers, c->stream);
  return launched(c);
}

int sbslpm_swe_displacement(sbslpm_ctx* c, const double* vx, const double* vy, const double* vz, double tau,
                            double* dpx, double* dpy, double* dpz, double* dxx, double* dxy, double* dxz) {
  if (!c || !vx || !vy || !vz || !dpx || !dpy || !dpz || !dxx || !dxy || !dxz)
    return fail(c, SBSLPM_E_ARG, "swe_displacement: null pointer");
  const double* v[3] = {vx, vy, vz};
  double* dp[3] = {dpx, dpy, dpz};
  double* dxp[3] = {dxx, dxy, dxz};
  launch_bve_displacement(c->dt, nullptr, nullptr, nullptr, v, nullptr, tau, 0.0, nullptr, dp, dxp, c->stream);
  return launched(c);
}

int sbslpm_swe_find_departure_points(sbslpm_ctx* c, const double* x, const double* y, const double* z,
                                     const double* const* vels, double dt, double* dep_x, double* dep_y,
                                     double* dep_z) {
  if (!c || !x || !y || !z || !vels || !dep_x || !dep_y || !dep_z)
    return fail(c, SBSLPM_E_ARG, "swe_find_departure_points: null pointer");
  for (int i = 0; i < 12; ++i)
    if (!vels[i]) return fail(c, SBSLPM_E_ARG, "swe_find_departure_points: null velocity");
  launch_bve_departure(c->dt, x, y, z, vels, dt, c->prm.rk4_combine, dep_x, dep_y, dep_z, c->stream);
  return launched(c);
}

int sbslpm_swe_departure_to_target(sbslpm_ctx* c, const double* dep_x, const double* dep_y, const double* dep_z,
                                   const double* vor, const double* div, const double* h, const double* tracers,
                                   int ntr, double* new_vor, double* new_div, double* new_h, double* new_tracers) {
  if (!c || !dep_x || !dep_y || !dep_z || !vor || !div || !h || !new_vor || !new_div || !new_h || ntr < 0 ||
      (ntr > 0 && (!tracers || !new_tracers)))
    return fail(c, SBSLPM_E_ARG, "swe_departure_to_target: bad argument");
  const double* f[3] = {vor, div, h};
  double* nf[3] = {new_vor, new_div, new_h};
  launch_swe_departure_to_target(c->dt, dep_x, dep_y, dep_z, f, tracers, ntr, nf, new_tracers, c->stream);
  return launched(c);
}

// ---- shallow-water step (swe_time_step.cpp) -----------------------------------

static int ensure_swe_step(sbslpm_ctx* c) {
  if (c->sws_arena) return SBSLPM_OK;
  const Tree& t = c->tree;
  const size_t NP = (size_t)t.npart, NQ = (size_t)t.P * t.n, T = (size_t)std::max(c->prm.tracer_count, 0);
  Arena A;
  size_t o_dp[3], o_dx[3], o_v[3][3], o_t[3][3], o_st[3], o_arr[3], o_dep[3];
  for (int i = 0; i < 3; ++i) o_dp[i] = A.take<double>(NP);
  for (int i = 0; i < 3; ++i) o_dx[i] = A.take<double>(NQ);
  const size_t o_eff = A.take<double>(NP), o_lap = A.take<double>(NP);
  for (int s = 0; s < 3; ++s)
    for (int i = 0; i < 3; ++i) o_v[s][i] = A.take<double>(NP);
  for (int s = 0; s < 3; ++s)
    for (int i = 0; i < 3; ++i) o_t[s][i] = A.take<double>(NP);
  for (int i = 0; i < 3; ++i) o_st[i] = A.take<double>(NP);
  const size_t o_da = A.take<double>(NP);
  for (int i = 0; i < 3; ++i) o_arr[i] = A.take<double>(NP);
  for (int i = 0; i < 3; ++i) o_dep[i] = A.take<double>(NP);
  const size_t o_ntr = A.take<double>(NP * T), o_qp = A.take<double>(NP), o_qx = A.take<double>(NQ),
               o_pp = A.take<double>(NP), o_px = A.take<double>(NQ), total = A.take<char>(256);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (hipMalloc(&c->sws_arena, total) != hipSuccess) {
    c->sws_arena = nullptr;
    return fail(c, SBSLPM_E_NOMEM, "shallow-water step workspace: out of device memory");
  }
  char* b = c->sws_arena;
  auto& w = c->sws;
  for (int i = 0; i < 3; ++i) {
    w.dpart[i] = (double*)(b + o_dp[i]);
    w.dprox[i] = (double*)(b + o_dx[i]);
    w.st[i] = (double*)(b + o_st[i]);
    w.arr[i] = (double*)(b + o_arr[i]);
    w.dep[i] = (double*)(b + o_dep[i]);
    for (int s = 0; s < 3; ++s) {
      w.v[s][i] = (double*)(b + o_v[s][i]);
      w.tend[s][i] = (double*)(b + o_t[s][i]);
    }
  }
  w.eff = (double*)(b + o_eff);
  w.lap = (double*)(b + o_lap);
  w.da = (double*)(b + o_da);
  w.ntr = (double*)(b + o_ntr);
  w.qp = (double*)(b + o_qp);
  w.qx = (double*)(b + o_qx);
  w.pp = (double*)(b + o_pp);
  w.px = (double*)(b + o_px);
  return SBSLPM_OK;
}

// Upward pass of area * f, the regularised-Laplacian pair sums over displaced
// sources, downward pass (swe_time_step.cpp:695-701): lap [Pl][m2] assigned.
static int laplacian_fmm(sbslpm_ctx* c, const double* area, const double* f, const double* const* dpart,
                         const double* const* dprox, double* lap) {
  const Tree& t = c->tree;
  auto& w = c->sws;
  launch_upward_fused(c->dt, area, f, w.qp, w.qx, c->stream);
  double* px1[1] = {w.px};
  launch_zero_rows(c->zero_rows, c->nzero_rows, t.n, 1, px1, c->stream);
  if (c->prm.precision != SBSLPM_FP64) {
    launch_source_records32(c->dt, w.qp, w.qx, dpart, dprox, c->rec32, c->stream);
    KernelTimer tm(c, SBSLPM_KERNEL_LAP_MIXED);
    launch_laplacian_interact_mixed(c->dt, c->tiles, c->ntiles, c->sets, c->rec32, w.pp, w.px, c->stream, true);
  } else {
    launch_source_records64(c->dt, w.qp, w.qx, dpart, dprox, true, c->rec64, c->stream);
    KernelTimer tm(c, SBSLPM_KERNEL_LAP_FP64);
    launch_laplacian_interact64(c->dt, c->tiles, c->ntiles, c->sets, c->rec64, w.pp, w.px, c->stream, true);
  }
  const double* pp1[1] = {w.pp};
  double* out[1] = {lap};
  launch_downward_pass(c->dt, 1, px1, pp1, out, c->stream, true);
  return launched(c);
}

// The context's shallow-water physics for the kernels.
static SwePhys swe_phys(const sbslpm_ctx* c) {
  const sbslpm_params& p = c->prm;
  SwePhys ph{};
  ph.topo = p.topography;
  ph.surface_correction = (p.swe_form & SBSLPM_SWE_FORM_SURFACE_CORRECTION) ? 1 : 0;
  ph.gravity = p.gravity > 0 ? p.gravity : 1.0;
  ph.mtn_h = p.mountain_height;
  ph.mtn_w = p.mountain_width > 0 ? p.mountain_width : 1.0;
  ph.mc[0] = cos(p.mountain_lat) * cos(p.mountain_lon);
  ph.mc[1] = cos(p.mountain_lat) * sin(p.mountain_lon);
  ph.mc[2] = sin(p.mountain_lat);
  return ph;
}

static bool swe_rule_ok(sbslpm_ctx* c) {
  return !(c->prm.topography && c->prm.swe_rule == SBSLPM_SWE_RULE_REFERENCE && c->tree.m2 != c->tree.n);
}

int sbslpm_laplacian(sbslpm_ctx* c, const double* f, double* lap) {
  if (!c || !f || !lap) return fail(c, SBSLPM_E_ARG, "laplacian: null pointer");
  launch_panel_laplacian(c->dt, f, lap, c->stream);
  return launched(c);
}

int sbslpm_laplacian_interactions(sbslpm_ctx* c, const double* dpx, const double* dpy, const double* dpz,
                                  const double* dxx, const double* dxy, const double* dxz, const double* q_part,
                                  const double* q_proxy, double* phi_part, double* phi_proxy) {
  if (!c || !q_part || !q_proxy || !phi_part || !phi_proxy)
    return fail(c, SBSLPM_E_ARG, "laplacian_interactions: null pointer");
  if ((dpx || dpy || dpz || dxx || dxy || dxz) && !(dpx && dpy && dpz && dxx && dxy && dxz))
    return fail(c, SBSLPM_E_ARG, "laplacian_interactions: partial displacement");
  const double* dp[3] = {dpx, dpy, dpz};
  const double* dx[3] = {dxx, dxy, dxz};
  if (c->prm.precision != SBSLPM_FP64) {
    launch_source_records32(c->dt, q_part, q_proxy, dpx ? dp : nullptr, dxx ? dx : nullptr, c->rec32, c->stream);
    launch_laplacian_interact_mixed(c->dt, c->tiles, c->ntiles, c->sets, c->rec32, phi_part, phi_proxy, c->stream,
                                    false);
    return launched(c);
  }
  launch_source_records64(c->dt, q_part, q_proxy, dpx ? dp : nullptr, dxx ? dx : nullptr, true, c->rec64, c->stream);
  launch_laplacian_interact64(c->dt, c->tiles, c->ntiles, c->sets, c->rec64, phi_part, phi_proxy, c->stream, false);
  return launched(c);
}

int sbslpm_apply_topography(sbslpm_ctx* c, const double* x, const double* y, const double* z, const double* dpx,
                            const double* dpy, const double* dpz, const double* dxx, const double* dxy,
                            const double* dxz, const double* height, double* eff) {
  if (!c || !x || !y || !z || !height || !eff) return fail(c, SBSLPM_E_ARG, "apply_topography: null pointer");
  if (!swe_rule_ok(c))
    return fail(c, SBSLPM_E_ARG, "apply_topography: the reference rule needs leaf_degree == interp_degree");
  if (!c->prm.topography) {
    launch_copy(eff, height, c->tree.npart, c->stream);
    return launched(c);
  }
  const double* dp[3] = {dpx, dpy, dpz};
  const double* dx[3] = {dxx, dxy, dxz};
  launch_topography(c->dt, x, y, z, dp, dx, height, c->prm.swe_rule, swe_phys(c), eff, c->stream);
  return launched(c);
}

int sbslpm_swe_tendency(sbslpm_ctx* c, const double* x, const double* y, const double* z, const double* vx,
                        const double* vy, const double* vz, const double* vor, const double* div, const double* h,
                        const double* lap, double* vor_t, double* div_t, double* h_t) {
  if (!c || !x || !y || !z || !vx || !vy || !vz || !vor || !div || !h || !lap || !vor_t || !div_t || !h_t)
    return fail(c, SBSLPM_E_ARG, "swe_tendency: null pointer");
  const double* v[3] = {vx, vy, vz};
  double* o[3] = {vor_t, div_t, h_t};
  launch_swe_tendency(c->dt, x, y, z, v, vor, div, h, lap, c->prm.omega, c->prm.swe_rule, swe_phys(c).gravity, o,
                      c->stream);
  return launched(c);
}

int sbslpm_swe_height_correction(sbslpm_ctx* c, const double* x, const double* y, const double* z, const double* vx,
                                 const double* vy, const double* vz, const double* h, const double* h_t, double dt,
                                 double* lap) {
  if (!c || !x || !y || !z || !vx || !vy || !vz || !h || !h_t || !lap)
    return fail(c, SBSLPM_E_ARG, "swe_height_correction: null pointer");
  const double* v[3] = {vx, vy, vz};
  launch_height_correction(c->dt, x, y, z, v, h, h_t, dt, swe_phys(c), lap, c->stream);
  return launched(c);
}

int sbslpm_unify_boundary_vals(sbslpm_ctx* c, const int* two_d_to_1d, const int* counts, int64_t n1d,
                               double* vals) {
  if (!c || !two_d_to_1d || !counts || !vals) return fail(c, SBSLPM_E_ARG, "unify_boundary_vals: null pointer");
  const int64_t n = (int64_t)c->tree.Pl * c->tree.m2;
  if (n1d < 1 || n1d > n) return fail(c, SBSLPM_E_ARG, "unify_boundary_vals: n1d must lie in [1, Pl*m2]");
  int rc = ensure_scratch(c, n1d);
  if (rc) return rc;
  launch_unify_boundary_vals(n, two_d_to_1d, counts, n1d, c->scratch, vals, c->stream);
  return launched(c);
}

static int swe_step_body(sbslpm_ctx* c, const double* x, const double* y, const double* z, const double* area,
                         double* vors, double* divs, double* height, double* vx, double* vy, double* vz,
                         double* tracers, double* height_lap, double* record = nullptr);

int sbslpm_swe_back_rk4_step(sbslpm_ctx* c, const double* x, const double* y, const double* z, const double* area,
                             double* vors, double* divs, double* height, double* vx, double* vy, double* vz,
                             double* tracers, double* height_lap, double time) {
  (void)time;
  if (!c || !x || !y || !z || !area || !vors || !divs || !height || !vx || !vy || !vz || !height_lap)
    return fail(c, SBSLPM_E_ARG, "swe_back_rk4_step: null pointer");
  const int ntr = c->prm.tracer_count;
  if (ntr > 0 && !tracers) return fail(c, SBSLPM_E_ARG, "swe_back_rk4_step: tracers required (tracer_count > 0)");
  {
    const void* outs[4] = {vx, vy, vz, height_lap};
    const void* ins[4] = {vors, divs, height, area};
    for (const void* o : outs)
      for (const void* i : ins)
        if (o == i) return fail(c, SBSLPM_E_ARG, "swe_back_rk4_step: velocity / height_lap must not alias the state");
    if (vors == divs || vors == height || divs == height)
      return fail(c, SBSLPM_E_ARG, "swe_back_rk4_step: vors, divs and height must be distinct arrays");
  }
  if (!swe_rule_ok(c))
    return fail(c, SBSLPM_E_ARG, "swe_back_rk4_step: the reference topography rule needs leaf_degree == interp_degree");
  int rc = ensure_swe(c);
  if (!rc) rc = ensure_swe_step(c);
  if (rc) return rc;
  // Graph capture needs a non-default stream; on the default stream the step runs eagerly.
  if (!(c->prm.flags & SBSLPM_FLAG_GRAPH) || !c->stream)
    return swe_step_body(c, x, y, z, area, vors, divs, height, vx, vy, vz, tracers, height_lap);
  const void* key[13] = {x,  y,  z,  area,    vors,       divs, height, vx, vy, vz, tracers, height_lap,
                         (const void*)(intptr_t)c->prm.precision};
  if (!c->swe_graph || std::memcmp(key, c->swe_graph_key, sizeof(key)) != 0) {
    if (c->swe_graph) {
      hipGraphExecDestroy(c->swe_graph);
      c->swe_graph = nullptr;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    rc = swe_step_body(c, x, y, z, area, vors, divs, height, vx, vy, vz, tracers, height_lap);
    c->capturing = false;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(c, SBSLPM_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->swe_graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
      c->swe_graph = nullptr;
      return fail(c, SBSLPM_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    std::memcpy(c->swe_graph_key, key, sizeof(key));
  }
  HIP_TRY(c, hipGraphLaunch(c->swe_graph, c->stream));
  return SBSLPM_OK;
}

// The kernels of one shallow-water step (eager or under capture); the
// workspaces exist already.  record (optional, [21][Pl][m2]): the stage
// velocities V0..V3 and the tendencies T0..T2 the reverse pass reads.
static int swe_step_body(sbslpm_ctx* c, const double* x, const double* y, const double* z, const double* area,
                         double* vors, double* divs, double* height, double* vx, double* vy, double* vz,
                         double* tracers, double* height_lap, double* record) {
  int rc = SBSLPM_OK;
  const int ntr = c->prm.tracer_count;
  const Tree& t = c->tree;
  auto& w = c->sws;
  const int rule = c->prm.swe_rule, topo = c->prm.topography;
  const SwePhys ph = swe_phys(c);
  const double dt = c->prm.delta_t;
  const double taus[3] = {0.5 * dt, 0.5 * dt, dt};
  const double* state[3] = {vors, divs, height};
  double* vel1[3] = {vx, vy, vz};
  const double* cdp[3] = {w.dpart[0], w.dpart[1], w.dpart[2]};
  const double* cdx[3] = {w.dprox[0], w.dprox[1], w.dprox[2]};
  // stage 1: undisplaced; the spectral Laplacian of the effective height
  const double* eff = height;
  if (topo) {
    launch_topography(c->dt, x, y, z, nullptr, nullptr, height, rule, ph, w.eff, c->stream);
    eff = w.eff;
  }
  if (c->prm.swe_form & SBSLPM_SWE_FORM_REG_LAPLACIAN) {
    // the operator of stages 2-3 at the undisplaced nodes: regularised-kernel
    // sums of area * eff, less 8/0.1 times the node's own height (or free surface)
    rc = laplacian_fmm(c, area, eff, nullptr, nullptr, height_lap);
    if (rc) return rc;
    launch_node_correction(t.npart, x, y, z, height, ph, height_lap, c->stream);
  } else {
    launch_panel_laplacian(c->dt, eff, height_lap, c->stream);
  }
  rc = sbslpm_swe_velocity(c, 1, area, vors, divs, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, vx, vy, vz);
  if (rc) return rc;
  launch_swe_tendency(c->dt, x, y, z, vel1, vors, divs, height, height_lap, c->prm.omega, rule, ph.gravity,
                      w.tend[0], c->stream);
  // stages 2..4
  for (int s = 1; s <= 3; ++s) {
    double* const* vprev = s == 1 ? vel1 : w.v[s - 2];
    const double* cvprev[3] = {vprev[0], vprev[1], vprev[2]};
    rc = sbslpm_swe_displacement(c, vprev[0], vprev[1], vprev[2], taus[s - 1], w.dpart[0], w.dpart[1], w.dpart[2],
                                 w.dprox[0], w.dprox[1], w.dprox[2]);
    if (rc) return rc;
    const double* tp[3] = {w.tend[s - 1][0], w.tend[s - 1][1], w.tend[s - 1][2]};
    const double tau_apply = (s == 3 && rule == SBSLPM_SWE_RULE_REFERENCE) ? 0.5 * dt : taus[s - 1];
    launch_apply_tendencies(t.npart, state, tp, area, tau_apply, w.st, w.da, c->stream);
    rc = sbslpm_swe_velocity(c, 1, w.da, w.st[0], w.st[1], w.dpart[0], w.dpart[1], w.dpart[2], w.dprox[0], w.dprox[1],
                             w.dprox[2], w.v[s - 1][0], w.v[s - 1][1], w.v[s - 1][2]);
    if (rc) return rc;
    if (s == 3) break;
    const double* effs = w.st[2];
    if (topo) {
      launch_topography(c->dt, x, y, z, cdp, cdx, w.st[2], rule, ph, w.eff, c->stream);
      effs = w.eff;
    }
    rc = laplacian_fmm(c, w.da, effs, cdp, cdx, w.lap);
    if (rc) return rc;
    launch_height_correction(c->dt, x, y, z, cvprev, height, w.tend[s - 1][2], taus[s - 1], ph, w.lap, c->stream);
    const double* cv[3] = {w.v[s - 1][0], w.v[s - 1][1], w.v[s - 1][2]};
    launch_swe_tendency(c->dt, x, y, z, cv, vors, divs, height, w.lap, c->prm.omega, rule, ph.gravity, w.tend[s],
                        c->stream);
  }
  // arrival values (state + dt k1), departure points, remap
  const double* t1[3] = {w.tend[0][0], w.tend[0][1], w.tend[0][2]};
  launch_apply_tendencies(t.npart, state, t1, nullptr, dt, w.arr, nullptr, c->stream);
  const double* vels[12] = {vx, vy, vz, w.v[0][0], w.v[0][1], w.v[0][2],
                            w.v[1][0], w.v[1][1], w.v[1][2], w.v[2][0], w.v[2][1], w.v[2][2]};
  launch_bve_departure(c->dt, x, y, z, vels, dt, c->prm.rk4_combine, w.dep[0], w.dep[1], w.dep[2], c->stream);
  const double* arr[3] = {w.arr[0], w.arr[1], w.arr[2]};
  double* out[3] = {vors, divs, height};
  if (record) {
    const double* src[21] = {vx, vy, vz};
    for (int s2 = 0; s2 < 3; ++s2)
      for (int q = 0; q < 3; ++q) {
        src[3 + 3 * s2 + q] = w.v[s2][q];
        src[12 + 3 * s2 + q] = w.tend[s2][q];
      }
    for (int i = 0; i < 21; ++i) launch_copy(record + (int64_t)i * t.npart, src[i], t.npart, c->stream);
  }
  launch_swe_departure_to_target(c->dt, w.dep[0], w.dep[1], w.dep[2], arr, tracers, ntr, out, w.ntr, c->stream);
  if (ntr > 0) launch_copy(tracers, w.ntr, t.npart * ntr, c->stream);
  return launched(c);
}

// ---- shallow-water step adjoint ---------------------------------------------------

static int ensure_adjoint(sbslpm_ctx* c);

static int ensure_swe_adjoint(sbslpm_ctx* c) {
  int rc = ensure_adjoint(c);
  if (!rc) rc = ensure_swe(c);
  if (!rc) rc = ensure_swe_step(c);
  if (rc || c->swa_arena) return rc;
  const Tree& t = c->tree;
  const size_t N = (size_t)t.npart, Pn = (size_t)t.P * t.n;
  Arena A;
  const size_t o_vb = A.take<double>(12 * N), o_tb = A.take<double>(9 * N), o_rb = A.take<double>(3 * N),
               o_lb = A.take<double>(N), o_qp = A.take<double>(3 * N), o_qx = A.take<double>(3 * Pn),
               o_q = A.take<double>(3 * N), o_dbp = A.take<double>(3 * N), o_dbx = A.take<double>(3 * Pn),
               o_hs = A.take<double>(N), o_hh = A.take<double>(N), o_st = A.take<double>(21 * N),
               o_tr = A.take<char>(t.U * swe_adj_target_record_bytes()), o_sr = A.take<SrcRec64>(t.U),
               total = A.take<char>(256);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  char* b = nullptr;
  if (hipMalloc(&b, total) != hipSuccess)
    return fail(c, SBSLPM_E_NOMEM, "shallow-water adjoint workspace: out of device memory");
  c->swa_arena = b;
  auto& a = c->swa;
  for (int i = 0; i < 12; ++i) a.vb[i] = (double*)(b + o_vb) + i * N;
  for (int i = 0; i < 9; ++i) a.tb[i] = (double*)(b + o_tb) + i * N;
  for (int i = 0; i < 3; ++i) {
    a.rb[i] = (double*)(b + o_rb) + i * N;
    a.dbp[i] = (double*)(b + o_dbp) + i * N;
    a.dbx[i] = (double*)(b + o_dbx) + i * Pn;
  }
  a.lapbar = (double*)(b + o_lb);
  a.qzp = (double*)(b + o_qp);
  a.qdp = a.qzp + N;
  a.qep = a.qdp + N;
  a.qzx = (double*)(b + o_qx);
  a.qdx = a.qzx + Pn;
  a.qex = a.qdx + Pn;
  a.qz = (double*)(b + o_q);
  a.qd = a.qz + N;
  a.qe = a.qd + N;
  a.hscat = (double*)(b + o_hs);
  a.hh = (double*)(b + o_hh);
  a.stages = (double*)(b + o_st);
  a.trec = b + o_tr;
  a.srec = (SrcRec64*)(b + o_sr);
  return SBSLPM_OK;
}

// Anterpolation of a particle cotangent with unit weights, the first step of
// every transposed FMM: adj_psip[f] = vbar[f] and the non-leaf rows of
// adj_psix[f] (no target-record kernel reads the leaf rows), as upward-pass
// charge sets, so that callers launch them together with their charges'
// upward passes (launch_upward_sets).
static void anterpolation_sets(sbslpm_ctx* c, const double* const* vbar, int ncomp, UpCharge* up) {
  for (int f = 0; f < ncomp; ++f) up[f] = {c->adj_ones, vbar[f], c->adj_psip[f], c->adj_psix[f]};
}

// Exact transpose of the shallow-water velocity evaluation (sbslpm_swe_velocity,
// variant 1, sources displaced by dpart/dprox or not) for the particle-velocity
// cotangent vbar: swa.qz, swa.qd = dL/d(area_s vor_s), dL/d(area_s div_s) at
// the particles; with displacements, swa.dbp/dbx = dL/d disp, assigned or (acc)
// added.  The pair sums run in FP32 on mixed-precision contexts.
static int swe_fmm_adjoint(sbslpm_ctx* c, const double* area_s, const double* vor_s, const double* div_s,
                           const double* const* dpart, const double* const* dprox, const double* const* vbar,
                           bool acc) {
  const Tree& t = c->tree;
  auto& a = c->swa;
  const bool pos = dpart != nullptr, fp32 = c->prm.precision != SBSLPM_FP64;
  {  // the cotangent's anterpolation and the charges' upward pass share their launches
    UpCharge up[5];
    anterpolation_sets(c, vbar, 3, up);
    up[3] = {area_s, vor_s, c->qpart, c->qproxy};
    up[4] = {area_s, div_s, c->qdpart, c->qdproxy};
    launch_upward_sets(c->dt, up, 5, c->stream);
  }
  const double* psix[3] = {c->adj_psix[0], c->adj_psix[1], c->adj_psix[2]};
  launch_swe_adj_target_records(c->dt, fp32, vbar, psix, a.trec, c->stream);
  launch_source_records_swe(c->dt, 1, c->qpart, c->qproxy, c->qdpart, c->qdproxy, dpart, dprox, c->recswe, c->stream);
  double* qx2[2] = {a.qzx, a.qdx};
  launch_zero_rows(c->adj_zero_rows, c->adj_nzero, t.n, 2, qx2, c->stream);
  if (pos && !acc) launch_zero_rows(c->adj_zero_rows, c->adj_nzero, t.n, 3, a.dbx, c->stream);
  launch_swe_adjoint(c->dt, fp32, c->adj_tiles, c->adj_ntiles, c->adj_targets, a.trec, c->recswe, c->prm.kernel_eps,
                     dpart, dprox, a.qzp, a.qzx, a.qdp, a.qdx, a.dbp, a.dbx, acc, c->stream, c->shape);
  // both charge cotangents in one downward pass (two components per launch)
  const double* qp2[2] = {a.qzp, a.qdp};
  double* q2[2] = {a.qz, a.qd};
  launch_downward_pass(c->dt, 2, qx2, qp2, q2, c->stream, true);
  return launched(c);
}

// Exact transpose of laplacian_fmm for the cotangent lbar of its output:
// swa.qe = dL/d(area_s eff) at the particles; with displacements swa.dbp/dbx
// = dL/d disp assigned.
static int lap_fmm_adjoint(sbslpm_ctx* c, const double* area_s, const double* eff, const double* const* dpart,
                           const double* const* dprox, const double* lbar) {
  const Tree& t = c->tree;
  auto& a = c->swa;
  auto& w = c->sws;
  const bool pos = dpart != nullptr, fp32 = c->prm.precision != SBSLPM_FP64;
  UpCharge up[2];
  anterpolation_sets(c, &lbar, 1, up);
  up[1] = {area_s, eff, w.qp, w.qx};
  launch_upward_sets(c->dt, up, 2, c->stream);  // the cotangent's anterpolation and the charge's upward pass
  launch_lap_adj_target_records(c->dt, fp32, lbar, c->adj_psix[0], a.trec, c->stream);
  launch_source_records64(c->dt, w.qp, w.qx, dpart, dprox, true, a.srec, c->stream);
  launch_zero_rows(c->adj_zero_rows, c->adj_nzero, t.n, 1, &a.qex, c->stream);
  if (pos) launch_zero_rows(c->adj_zero_rows, c->adj_nzero, t.n, 3, a.dbx, c->stream);
  launch_laplacian_adjoint(c->dt, fp32, c->adj_tiles, c->adj_ntiles, c->adj_targets, a.trec, a.srec, dpart, dprox,
                           a.qep, a.qex, a.dbp, a.dbx, false, c->stream, c->shape);
  const double* qep[1] = {a.qep};
  launch_downward_pass(c->dt, 1, &a.qex, qep, &a.qe, c->stream, true);
  return launched(c);
}

// Reverse pass of one shallow-water step from its record (stage velocities
// V0..V3 = rec[0..11], tendencies T0..T2 = rec[12..20], each [Pl][m2]):
// sbar (vor, div, h before the step) assigned from nbar (after the step).
static int swe_step_reverse(sbslpm_ctx* c, const double* x, const double* y, const double* z, const double* area,
                            const double* vors, const double* divs, const double* height, const double* rec,
                            const double* const* nbar, double* const* sbar) {
  const Tree& t = c->tree;
  const int64_t N = t.npart;
  auto& w = c->sws;
  auto& a = c->swa;
  const SwePhys ph = swe_phys(c);
  const int rule = c->prm.swe_rule, topo = c->prm.topography;
  const bool reg = (c->prm.swe_form & SBSLPM_SWE_FORM_REG_LAPLACIAN) != 0;
  const double dt = c->prm.delta_t, om = c->prm.omega;
  const double taus[3] = {0.5 * dt, 0.5 * dt, dt};
  const double* V[12];
  const double* T[9];
  for (int i = 0; i < 12; ++i) V[i] = rec + i * N;
  for (int i = 0; i < 9; ++i) T[i] = rec + (12 + i) * N;
  const double* state[3] = {vors, divs, height};
  const double* csbar[3] = {sbar[0], sbar[1], sbar[2]};
  {
    double* zero[6] = {sbar[0], sbar[1], sbar[2], a.vb[0], a.tb[0], a.rb[0]};
    const int64_t count[6] = {N, N, N, 12 * N, 9 * N, 3 * N};
    launch_fill_zero_sets(zero, count, 6, c->stream);
  }
  // arrival values R = S0 + dt T0, departure search and remap
  launch_apply_tendencies(N, state, T, nullptr, dt, w.arr, nullptr, c->stream);
  const double* arr[3] = {w.arr[0], w.arr[1], w.arr[2]};
  launch_departure_adjoint3(c->dt, x, y, z, V, arr, nbar, dt, c->prm.rk4_combine, a.vb, a.rb, c->stream);
  const double* crb[3] = {a.rb[0], a.rb[1], a.rb[2]};
  const double* ctb0[3] = {a.tb[0], a.tb[1], a.tb[2]};
  launch_apply_two(N, crb, csbar, 1.0, sbar, ctb0, dt, a.tb, c->stream);
  const double* cdp[3] = {w.dpart[0], w.dpart[1], w.dpart[2]};
  const double* cdx[3] = {w.dprox[0], w.dprox[1], w.dprox[2]};
  int rc = SBSLPM_OK;
  for (int s = 3; s >= 1 && rc == SBSLPM_OK; --s) {
    const double* const* vprev = V + 3 * (s - 1);
    const double* const* tprev = T + 3 * (s - 1);
    const double tau = taus[s - 1];
    const double taua = (s == 3 && rule == SBSLPM_SWE_RULE_REFERENCE) ? 0.5 * dt : tau;
    // the stage's displacement, displaced state and areas, as the forward
    rc = sbslpm_swe_displacement(c, vprev[0], vprev[1], vprev[2], tau, w.dpart[0], w.dpart[1], w.dpart[2], w.dprox[0],
                                 w.dprox[1], w.dprox[2]);
    if (rc) break;
    launch_apply_tendencies(N, state, tprev, area, taua, w.st, w.da, c->stream);
    if (s < 3) {
      // tendencies of this stage
      launch_swe_tendency_adjoint(c->dt, x, y, z, V + 3 * s, vors, divs, height, a.tb + 3 * s, om, rule, ph.gravity,
                                  a.vb + 3 * s, a.lapbar, sbar, c->stream);
      // height correction: lap -= 80 BLI[h + tau T_h](P(x + tau vprev)) (+ h_s)
      launch_lincomb_zero(N, height, tau, tprev[2], a.hh, a.hscat, c->stream);
      launch_height_correction_adjoint(c->dt, x, y, z, vprev, a.hh, tau, ph, a.lapbar, a.hscat, a.vb + 3 * (s - 1),
                                       c->stream);
      launch_add2(N, a.hscat, sbar[2], tau, a.tb[3 * (s - 1) + 2], c->stream);
      // regularised-Laplacian sums of area_s * eff at displaced sources
      const double* eff = w.st[2];
      if (topo) {
        launch_topography(c->dt, x, y, z, cdp, cdx, w.st[2], rule, ph, w.eff, c->stream);
        eff = w.eff;
      }
      rc = lap_fmm_adjoint(c, w.da, eff, cdp, cdx, a.lapbar);
      if (rc) break;
    }
    const double* vb[3] = {a.vb[3 * s], a.vb[3 * s + 1], a.vb[3 * s + 2]};
    rc = swe_fmm_adjoint(c, w.da, w.st[0], w.st[1], cdp, cdx, vb, s < 3);
    if (rc) break;
    launch_swe_stage_adjoint(c->dt, x, y, z, state, tprev, area, taua, a.qz, a.qd, s < 3 ? a.qe : nullptr,
                             (s < 3 && topo) ? cdp : nullptr, (s < 3 && topo) ? cdx : nullptr, rule, ph, sbar,
                             a.tb + 3 * (s - 1), a.dbp, a.dbx, c->stream);
    const double* dbx[3] = {a.dbx[0], a.dbx[1], a.dbx[2]};
    launch_disp_interp_adjoint(c->dt, dbx, a.dbp, c->stream);
    const double* cvb[3] = {a.vb[3 * (s - 1)], a.vb[3 * (s - 1) + 1], a.vb[3 * (s - 1) + 2]};
    const double* cdbp[3] = {a.dbp[0], a.dbp[1], a.dbp[2]};
    launch_apply_tendencies(N, cvb, cdbp, nullptr, tau, a.vb + 3 * (s - 1), nullptr, c->stream);
  }
  if (rc) return rc;
  // stage 1: tendencies, surface Laplacian of h (+ h_s), undisplaced velocity
  launch_swe_tendency_adjoint(c->dt, x, y, z, V, vors, divs, height, a.tb, om, rule, ph.gravity, a.vb, a.lapbar, sbar,
                              c->stream);
  if (reg) {
    // lap -= 8/0.1 (h + h_s(x)) at the nodes
    launch_lincomb(N, sbar[2], -8.0 / 0.1, a.lapbar, sbar[2], c->stream);
    const double* eff = height;
    if (topo) {
      launch_topography(c->dt, x, y, z, nullptr, nullptr, height, rule, ph, w.eff, c->stream);
      eff = w.eff;
    }
    rc = lap_fmm_adjoint(c, area, eff, nullptr, nullptr, a.lapbar);
    if (rc) return rc;
  } else {
    launch_panel_laplacian_adjoint(c->dt, a.lapbar, sbar[2], c->stream);
  }
  const double* vb0[3] = {a.vb[0], a.vb[1], a.vb[2]};
  rc = swe_fmm_adjoint(c, area, vors, divs, nullptr, nullptr, vb0, false);
  if (rc) return rc;
  launch_swe_stage_adjoint(c->dt, x, y, z, state, nullptr, area, 0.0, a.qz, a.qd, reg ? a.qe : nullptr, nullptr,
                           nullptr, rule, ph, sbar, nullptr, nullptr, nullptr, c->stream);
  return launched(c);
}

// Checks shared by the shallow-water record and adjoint entry points.
static int swe_adjoint_args(sbslpm_ctx* c, const char* what) {
  if (c->prm.tracer_count != 0)
    return fail(c, SBSLPM_E_ARG, std::string(what) + ": needs tracer_count == 0 (tracers are not differentiated)");
  if (!swe_rule_ok(c))
    return fail(c, SBSLPM_E_ARG, std::string(what) + ": the reference topography rule needs leaf_degree == interp_degree");
  return ensure_swe_adjoint(c);
}

extern "C" {

int sbslpm_swe_back_rk4_step_record(sbslpm_ctx* c, const double* x, const double* y, const double* z,
                                    const double* area, const double* vors, const double* divs, const double* height,
                                    double* new_vors, double* new_divs, double* new_height, double* stages) {
  if (!c || !x || !y || !z || !area || !vors || !divs || !height || !new_vors || !new_divs || !new_height || !stages)
    return fail(c, SBSLPM_E_ARG, "swe_back_rk4_step_record: null pointer");
  const double* ins[4] = {vors, divs, height, area};
  const double* outs[4] = {new_vors, new_divs, new_height, stages};
  for (const double* o : outs)
    for (const double* i : ins)
      if (o == i) return fail(c, SBSLPM_E_ARG, "swe_back_rk4_step_record: outputs must not alias the inputs");
  int rc = swe_adjoint_args(c, "swe_back_rk4_step_record");
  if (rc) return rc;
  const int64_t N = c->tree.npart;
  launch_copy(new_vors, vors, N, c->stream);
  launch_copy(new_divs, divs, N, c->stream);
  launch_copy(new_height, height, N, c->stream);
  return swe_step_body(c, x, y, z, area, new_vors, new_divs, new_height, stages, stages + N, stages + 2 * N, nullptr,
                       c->sws.lap, stages);
}

int sbslpm_swe_back_rk4_step_adjoint_stages(sbslpm_ctx* c, const double* x, const double* y, const double* z,
                                            const double* area, const double* vors, const double* divs,
                                            const double* height, const double* stages, const double* new_vors_bar,
                                            const double* new_divs_bar, const double* new_height_bar,
                                            double* vors_bar, double* divs_bar, double* height_bar) {
  if (!c || !x || !y || !z || !area || !vors || !divs || !height || !stages || !new_vors_bar || !new_divs_bar ||
      !new_height_bar || !vors_bar || !divs_bar || !height_bar)
    return fail(c, SBSLPM_E_ARG, "swe_back_rk4_step_adjoint_stages: null pointer");
  const double* ins[8] = {vors, divs, height, area, stages, new_vors_bar, new_divs_bar, new_height_bar};
  const double* outs[3] = {vors_bar, divs_bar, height_bar};
  for (const double* o : outs)
    for (const double* i : ins)
      if (o == i) return fail(c, SBSLPM_E_ARG, "swe_back_rk4_step_adjoint_stages: cotangents must not alias the inputs");
  int rc = swe_adjoint_args(c, "swe_back_rk4_step_adjoint_stages");
  if (rc) return rc;
  const double* nb[3] = {new_vors_bar, new_divs_bar, new_height_bar};
  double* sb[3] = {vors_bar, divs_bar, height_bar};
  return swe_step_reverse(c, x, y, z, area, vors, divs, height, stages, nb, sb);
}

int sbslpm_swe_back_rk4_step_adjoint(sbslpm_ctx* c, const double* x, const double* y, const double* z,
                                     const double* area, const double* vors, const double* divs, const double* height,
                                     const double* new_vors_bar, const double* new_divs_bar,
                                     const double* new_height_bar, double* vors_bar, double* divs_bar,
                                     double* height_bar) {
  if (!c || !x || !y || !z || !area || !vors || !divs || !height || !new_vors_bar || !new_divs_bar ||
      !new_height_bar || !vors_bar || !divs_bar || !height_bar)
    return fail(c, SBSLPM_E_ARG, "swe_back_rk4_step_adjoint: null pointer");
  const double* ins[7] = {vors, divs, height, area, new_vors_bar, new_divs_bar, new_height_bar};
  const double* outs[3] = {vors_bar, divs_bar, height_bar};
  for (const double* o : outs)
    for (const double* i : ins)
      if (o == i) return fail(c, SBSLPM_E_ARG, "swe_back_rk4_step_adjoint: cotangents must not alias the inputs");
  int rc = swe_adjoint_args(c, "swe_back_rk4_step_adjoint");
  if (rc) return rc;
  // forward recompute into the workspace's record; the stepped state itself is
  // not needed (written to the arrival-value buffers, then dropped)
  const int64_t N = c->tree.npart;
  auto& w = c->sws;
  double* tmp[3] = {c->swa.rb[0], c->swa.rb[1], c->swa.rb[2]};
  for (int q = 0; q < 3; ++q) launch_copy(tmp[q], (q == 0 ? vors : q == 1 ? divs : height), N, c->stream);
  rc = swe_step_body(c, x, y, z, area, tmp[0], tmp[1], tmp[2], c->swa.stages, c->swa.stages + N,
                     c->swa.stages + 2 * N, nullptr, w.lap, c->swa.stages);
  if (rc) return rc;
  const double* nb[3] = {new_vors_bar, new_divs_bar, new_height_bar};
  double* sb[3] = {vors_bar, divs_bar, height_bar};
  return swe_step_reverse(c, x, y, z, area, vors, divs, height, c->swa.stages, nb, sb);
}

}  // extern "C"

// ---- irregular particle tree (lat-lon path) ----------------------------------

// ---- autotuner -----------------------------------------------------------------

int sbslpm_autotune(sbslpm_ctx* c, const double* area, const double* vor, double* report, int* ncand) {
  if (!c || !area || !vor) return fail(c, SBSLPM_E_ARG, "autotune: null argument");
  if (c->capturing) return fail(c, SBSLPM_E_ARG, "autotune: not allowed during graph capture");
  const Work& w = c->w1;
  // the pair kernel's inputs: the context's records of (area, vor), undisplaced
  launch_upward_fused(c->dt, area, vor, w.qpart, w.qproxy, c->stream);
  const int prec = c->prm.precision;
  if (prec == SBSLPM_MIXED)
    launch_source_records_mfma(c->dt, w.qpart, w.qproxy, nullptr, nullptr, w.recmf, c->stream);
  else if (prec == SBSLPM_MIXED_VALU)
    launch_source_records32(c->dt, w.qpart, w.qproxy, nullptr, nullptr, w.rec32, c->stream);
  else
    launch_source_records64(c->dt, w.qpart, w.qproxy, nullptr, nullptr, true, w.rec64, c->stream);
  auto launch = [&](PairShape sh) {
    if (prec == SBSLPM_MIXED)
      launch_bve_interact_mfma(c->dt, c->tiles, c->ntiles, c->sets, w.recmf, c->prm.kernel_eps, w.phipart, w.phiproxy,
                               c->stream, true, 1, sh);
    else if (prec == SBSLPM_MIXED_VALU)
      launch_bve_interact_mixed(c->dt, c->tiles, c->ntiles, c->sets, w.rec32, c->prm.kernel_eps, w.phipart,
                                w.phiproxy, c->stream, true, 1, sh);
    else
      launch_bve_interact64(c->dt, c->tiles, c->ntiles, c->sets, w.rec64, c->prm.kernel_eps, w.phipart, w.phiproxy,
                            c->stream, true, 1, sh);
  };
  hipEvent_t e0, e1;
  HIP_TRY(c, hipEventCreate(&e0));
  HIP_TRY(c, hipEventCreate(&e1));
  const int waves[3] = {1, 2, 4};
  PairShape best = c->shape;
  float best_ms = 1e30f;
  int k = 0;
  for (int xcd = 0; xcd < 3; ++xcd)
    for (int wi = 0; wi < 3; ++wi) {
      const PairShape sh = full_list_shape(c->tree, waves[wi], xcd, prec == SBSLPM_MIXED);
      launch(sh);  // warm-up
      hipEventRecord(e0, c->stream);
      for (int r = 0; r < 3; ++r) launch(sh);
      hipEventRecord(e1, c->stream);
      if (hipEventSynchronize(e1) != hipSuccess) break;
      float ms = 0;
      hipEventElapsedTime(&ms, e0, e1);
      ms /= 3;
      if (report) {
        report[3 * k] = sh.waves;
        report[3 * k + 1] = sh.xcd;
        report[3 * k + 2] = ms;
      }
      ++k;
      if (ms < best_ms) {
        best_ms = ms;
        best = sh;
      }
    }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (ncand) *ncand = k;
  int rc = launched(c);
  if (rc) return rc;
  if (best.waves != c->shape.waves || best.xcd != c->shape.xcd) {
    c->shape = best;
    // captured graphs hold the old launch shape
    if (c->graph) {
      hipGraphExecDestroy(c->graph);
      c->graph = nullptr;
    }
    if (c->ens_graph) {
      hipGraphExecDestroy(c->ens_graph);
      c->ens_graph = nullptr;
    }
  }
  return SBSLPM_OK;
}

int sbslpm_set_pair_shape(sbslpm_ctx* c, int waves, int xcd) {
  if (!c || !(waves == 1 || waves == 2 || waves == 4) || !(xcd == 0 || xcd == 1 || xcd == 2))
    return fail(c, SBSLPM_E_ARG, "set_pair_shape: waves must be 1, 2 or 4 and xcd 0, 1 or 2");
  if (c->capturing) return fail(c, SBSLPM_E_ARG, "set_pair_shape: not allowed during graph capture");
  if (waves != c->shape.waves || xcd != c->shape.xcd) {
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->shape = full_list_shape(c->tree, waves, xcd, c->prm.precision == SBSLPM_MIXED);
    if (c->graph) {
      hipGraphExecDestroy(c->graph);
      c->graph = nullptr;
    }
    if (c->ens_graph) {
      hipGraphExecDestroy(c->ens_graph);
      c->ens_graph = nullptr;
    }
  }
  return SBSLPM_OK;
}

int sbslpm_get_pair_shape(const sbslpm_ctx* c, int* waves, int* xcd) {
  if (!c || !waves || !xcd) return SBSLPM_E_ARG;
  *waves = c->shape.waves;
  *xcd = c->shape.xcd;
  return SBSLPM_OK;
}

int sbslpm_irreg_create(sbslpm_ctx* c, int64_t n, const double* x, const double* y, const double* z, int leaf_size,
                        int interp_degree, sbslpm_irreg** out) {
  if (!c || !out || !x || !y || !z) return fail(c, SBSLPM_E_ARG, "irreg_create: null argument");
  *out = nullptr;
  sbslpm_irreg* h = new (std::nothrow) sbslpm_irreg();
  if (!h) return fail(c, SBSLPM_E_NOMEM, "irreg_create: out of host memory");
  std::string msg = build_irreg_tree(x, y, z, n, leaf_size, interp_degree, c->prm.theta, h->tree);
  if (!msg.empty()) {
    delete h;
    return fail(c, SBSLPM_E_ARG, "irreg_create: " + msg);
  }
  h->ctx = c;
  const IrregTree& t = h->tree;
  const size_t N = (size_t)t.N, Pn = (size_t)t.P * t.n, U = (size_t)t.U;
  std::vector<int> count(t.P, 0), leaves;
  for (int p = 0; p < t.P; ++p) {
    count[p] = t.pn[p].point_count;
    if (t.pn[p].is_leaf) leaves.push_back(p);
  }
  // Each leaf's ancestors with the composite matrices that evaluate them at the
  // leaf's nodes (the position bit of each step up is the upper-half flag of
  // the child: quadrants (hi,hi), (lo,hi), (lo,lo), (hi,lo) in (xi, eta)).
  std::vector<int> anc_first(1, 0), anc;
  int maxgap = 0;
  for (int p : leaves) {
    int node = p, px = 0, py = 0;
    for (int g = 1; t.pn[node].parent >= 0; ++g) {
      const int par = t.pn[node].parent;
      int q = 0;
      while (q < 4 && t.pn[par].child[q] != node) ++q;
      if (q == 4) {
        delete h;
        return fail(c, SBSLPM_E_ARG, "irreg_create: a panel is not among its parent's children");
      }
      px |= (q == 0 || q == 3) << (g - 1);
      py |= (q <= 1) << (g - 1);
      anc.insert(anc.end(), {par, (1 << g) - 2 + px, (1 << g) - 2 + py});
      maxgap = std::max(maxgap, g);
      node = par;
    }
    anc_first.push_back((int)anc.size() / 3);
  }
  // The table holds 2^(maxgap+1) - 2 matrices; past a depth no tree of a
  // realistic point set reaches (clustered points could), the level-by-level
  // lists stay in use instead.
  constexpr int kMaxCompositeGap = 14;
  std::vector<double> anc_tab;
  if (maxgap <= kMaxCompositeGap)
    composite_l2l(t.degree, t.cheb.data(), t.cheb.data(), maxgap, anc_tab);
  else
    anc.clear();
  Arena A;
  const size_t o_pan = A.take<DevPanel>(t.P), o_px = A.take<double>(3 * U), o_cheb = A.take<double>(t.dp1),
               o_hi = A.take<double>(t.dp1 * t.dp1), o_lo = A.take<double>(t.dp1 * t.dp1), o_perm = A.take<int>(N),
               o_lf = A.take<int>(t.P), o_lc = A.take<int>(t.P), o_pl = A.take<int>(N),
               o_lv = A.take<int>(leaves.size()), o_nl = A.take<int>(t.nonleaf.size()),
               o_tl = A.take<TargetTile>(t.tiles.size()), o_st = A.take<SourceSet>(t.sets.size()),
               o_zr = A.take<int>(t.zero_rows.size()), o_q = A.take<double>(2 * N), o_qx = A.take<double>(2 * Pn),
               o_part = A.take<double>(3 * N), o_prox = A.take<double>(3 * Pn), o_rs = A.take<SrcRecSwe>(U),
               o_r64 = A.take<SrcRec64>(U), o_w = A.take<double>((size_t)N * 2 * t.dp1),
               o_af = A.take<int>(anc_first.size()), o_anc = A.take<int>(std::max<size_t>(anc.size(), 1)),
               o_atab = A.take<double>(std::max<size_t>(anc_tab.size(), 1)), total = A.take<char>(256);
  char* b = nullptr;
  if (hipMalloc(&b, total) != hipSuccess) {
    delete h;
    return fail(c, SBSLPM_E_NOMEM, "irreg_create: out of device memory");
  }
  h->arena = b;
  double* pxyz = (double*)(b + o_px);
  int rc = upload(c, b + o_pan, t.dev.data(), t.P * sizeof(DevPanel));
  if (!rc) rc = upload(c, pxyz, t.px.data(), U * sizeof(double));
  if (!rc) rc = upload(c, pxyz + U, t.py.data(), U * sizeof(double));
  if (!rc) rc = upload(c, pxyz + 2 * U, t.pz.data(), U * sizeof(double));
  if (!rc) rc = upload(c, b + o_cheb, t.cheb.data(), t.dp1 * sizeof(double));
  if (!rc) rc = upload(c, b + o_hi, t.m2m_hi.data(), t.m2m_hi.size() * sizeof(double));
  if (!rc) rc = upload(c, b + o_lo, t.m2m_lo.data(), t.m2m_lo.size() * sizeof(double));
  if (!rc) rc = upload(c, b + o_perm, t.perm.data(), N * sizeof(int));
  if (!rc) rc = upload(c, b + o_lf, t.leaf_first.data(), t.P * sizeof(int));
  if (!rc) rc = upload(c, b + o_lc, count.data(), t.P * sizeof(int));
  if (!rc) rc = upload(c, b + o_pl, t.point_leaf.data(), N * sizeof(int));
  if (!rc) rc = upload(c, b + o_lv, leaves.data(), leaves.size() * sizeof(int));
  if (!rc) rc = upload(c, b + o_nl, t.nonleaf.data(), t.nonleaf.size() * sizeof(int));
  if (!rc) rc = upload(c, b + o_af, anc_first.data(), anc_first.size() * sizeof(int));
  if (!rc) rc = upload(c, b + o_anc, anc.data(), anc.size() * sizeof(int));
  if (!rc) rc = upload(c, b + o_atab, anc_tab.data(), anc_tab.size() * sizeof(double));
  // tiles heaviest first (sum of their source-set sizes): the pair kernels take one tile per
  // workgroup in dispatch order, so the long tiles start first and the short ones fill in
  std::vector<TargetTile> tiles_by_cost = t.tiles;
  {
    auto cost = [&](const TargetTile& tl) {
      int64_t w = 0;
      for (int e = tl.src_begin; e < tl.src_end; ++e) w += t.sets[e].count;
      return w;
    };
    std::vector<int64_t> w(tiles_by_cost.size());
    std::vector<int> order(tiles_by_cost.size());
    for (size_t i = 0; i < order.size(); ++i) {
      order[i] = (int)i;
      w[i] = cost(t.tiles[i]);
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b2) { return w[a] > w[b2]; });
    for (size_t i = 0; i < order.size(); ++i) tiles_by_cost[i] = t.tiles[order[i]];
  }
  if (!rc) rc = upload(c, b + o_tl, tiles_by_cost.data(), tiles_by_cost.size() * sizeof(TargetTile));
  if (!rc) rc = upload(c, b + o_st, t.sets.data(), t.sets.size() * sizeof(SourceSet));
  if (!rc) rc = upload(c, b + o_zr, t.zero_rows.data(), t.zero_rows.size() * sizeof(int));
  if (rc) {
    hipFree(b);
    delete h;
    return rc;
  }
  DevTree& d = h->dt;
  d = c->dt;  // error word; the rest below
  d.panels = (const DevPanel*)(b + o_pan);
  d.px = pxyz; d.py = pxyz + U; d.pz = pxyz + 2 * U;
  d.cheb_d = d.cheb_m = (const double*)(b + o_cheb);
  d.m2m_hi = (const double*)(b + o_hi);
  d.m2m_lo = (const double*)(b + o_lo);
  d.m2m_hiT = d.m2m_loT = nullptr;
  d.p2m = d.p2mT = nullptr;
  d.anc_l2l = d.l2p_hi = d.l2p_lo = d.l2pT_hi = d.l2pT_lo = nullptr;
  d.proxy_leaf = nullptr; d.proxy_xieta = nullptr; d.proxy_order = nullptr;
  d.P = t.P; d.Pl = (int)leaves.size(); d.offset = t.P;  // every proxy row may be an interaction source
  d.n = t.n; d.dp1 = t.dp1; d.m = 0; d.m2 = 0; d.levels = t.levels;
  d.npart = t.N; d.U = t.U; d.m2p = 0; d.np16 = 0; d.U16 = 0;
  h->perm = (int*)(b + o_perm);
  h->leaf_first = (int*)(b + o_lf);
  h->leaf_count = (int*)(b + o_lc);
  h->point_leaf = (int*)(b + o_pl);
  h->leaves = (int*)(b + o_lv);
  h->nleaves = (int)leaves.size();
  h->nonleaf = (int*)(b + o_nl);
  h->anc_first = (int*)(b + o_af);
  h->anc = (int*)(b + o_anc);
  h->anc_tab = anc_tab.empty() ? nullptr : (double*)(b + o_atab);
  h->tiles = (TargetTile*)(b + o_tl);
  h->ntiles = (int)t.tiles.size();
  h->sets = (SourceSet*)(b + o_st);
  h->zero_rows = (int*)(b + o_zr);
  h->nzero = (int)t.zero_rows.size();
  for (int k = 0; k < 2; ++k) {
    h->q[k] = (double*)(b + o_q) + k * N;
    h->qx[k] = (double*)(b + o_qx) + k * Pn;
  }
  for (int f = 0; f < 3; ++f) {
    h->part[f] = (double*)(b + o_part) + f * N;
    h->prox[f] = (double*)(b + o_prox) + f * Pn;
  }
  h->recswe = (SrcRecSwe*)(b + o_rs);
  h->rec64 = (SrcRec64*)(b + o_r64);
  h->w = (double*)(b + o_w);
  launch_point_weights(h->dt, h->point_leaf, h->w, c->stream);
  if (int lrc = launched(c)) {
    sbslpm_irreg_destroy(h);
    return lrc;
  }
  *out = h;
  return SBSLPM_OK;
}

void sbslpm_irreg_destroy(sbslpm_irreg* h) {
  if (!h) return;
  if (h->ctx && h->ctx->stream) hipStreamSynchronize(h->ctx->stream);
  if (h->arena) hipFree(h->arena);
  if (h->ll_ws) hipFree(h->ll_ws);
  delete h;
}

int sbslpm_irreg_sizes(const sbslpm_irreg* h, int64_t* s) {
  if (!h || !s) return SBSLPM_E_ARG;
  const IrregTree& t = h->tree;
  s[0] = t.P; s[1] = t.levels; s[2] = t.n; s[3] = t.N; s[4] = t.interaction_count(); s[5] = h->ntiles;
  return SBSLPM_OK;
}

}  // extern "C"

namespace {

// Proxy charges of every panel from point values: P2M at the leaves, then M2M
// bottom-up over the non-leaf panels (base_pots_ll + child_to_parent_ll).
void irreg_upward(sbslpm_irreg* h, const double* area, const double* pots, double* q_sorted, double* proxy) {
  const IrregTree& t = h->tree;
  hipStream_t s = h->ctx->stream;
  launch_gather_charges(t.N, h->perm, area, pots, q_sorted, s);
  launch_p2m_ll(h->dt, h->leaves, h->nleaves, h->leaf_first, h->leaf_count, h->w, 1, q_sorted, 0, proxy, 0, s);
  for (int l = t.levels; l >= 0; --l) {
    const int a = t.level_nonleaf[l], b = t.level_nonleaf[l + 1];
    launch_m2m_list(h->dt, h->nonleaf + a, b - a, proxy, s);
  }
}

// The same for vorticity and divergence together, into h->q[0..1] and h->qx[0..1]
// (each pair N and P*n apart): one P2M and one M2M launch per level for both.
void irreg_upward2(sbslpm_irreg* h, const double* area, const double* vors, const double* divs) {
  const IrregTree& t = h->tree;
  hipStream_t s = h->ctx->stream;
  launch_gather_charges(t.N, h->perm, area, vors, h->q[0], s);
  launch_gather_charges(t.N, h->perm, area, divs, h->q[1], s);
  launch_p2m_ll(h->dt, h->leaves, h->nleaves, h->leaf_first, h->leaf_count, h->w, 2, h->q[0], h->q[1] - h->q[0],
                h->qx[0], h->qx[1] - h->qx[0], s);
  for (int l = t.levels; l >= 0; --l) {
    const int a = t.level_nonleaf[l], b = t.level_nonleaf[l + 1];
    launch_m2m_list(h->dt, h->nonleaf + a, b - a, h->qx[0], s, 2);
  }
}

// L2L top-down over the non-leaf panels (accumulating into the children), then
// out_sorted = L2P(leaf rows) + part_sorted (parent_to_child_3_ll +
// child_panel_interp_3_ll).  With `fused` the sweep is one launch that forms
// the leaf rows from their ancestors' rows (equal in exact arithmetic) and
// leaves the non-leaf rows as they are.
void irreg_downward(sbslpm_irreg* h, double* const* proxy, const double* const* part_sorted, double* const* out_sorted,
                    bool fused = false) {
  const IrregTree& t = h->tree;
  hipStream_t s = h->ctx->stream;
  if (!fused || !launch_leaf_rows(h->dt, h->leaves, h->nleaves, h->anc_first, h->anc, h->anc_tab, 3, proxy, s))
    for (int l = 0; l <= t.levels; ++l) {
      const int a = t.level_nonleaf[l], b = t.level_nonleaf[l + 1];
      launch_l2l_list(h->dt, h->nonleaf + a, b - a, 3, proxy, s);
    }
  const double* cp[3] = {proxy[0], proxy[1], proxy[2]};
  launch_l2p_ll(h->dt, h->point_leaf, h->w, 3, cp, part_sorted, out_sorted, s);
}

// Pair sums of the lat-lon shallow-water kernel into the workspace (assigned).
void irreg_swe_pairs(sbslpm_irreg* h, const double* qz, const double* qzx, const double* qd, const double* qdx) {
  const IrregTree& t = h->tree;
  hipStream_t s = h->ctx->stream;
  launch_source_records_swe(h->dt, 3, qz, qzx, qd, qdx, nullptr, nullptr, h->recswe, s);
  launch_zero_rows(h->zero_rows, h->nzero, t.n, 3, h->prox, s);
  launch_swe_interact64(h->dt, 3, h->tiles, h->ntiles, h->sets, h->recswe, h->ctx->prm.kernel_eps, h->part, h->prox,
                        s, true, true);
}

// The lat-lon shallow-water velocity at the points in the caller's order (vel assigned):
// upward passes of both charges, pair sums, downward pass.
void irreg_swe_vel(sbslpm_irreg* h, const double* area, const double* vors, const double* divs, double* const* vel) {
  irreg_upward2(h, area, vors, divs);
  irreg_swe_pairs(h, h->q[0], h->qx[0], h->q[1], h->qx[1]);
  const double* cp[3] = {h->part[0], h->part[1], h->part[2]};
  irreg_downward(h, h->prox, cp, h->part, true);
  launch_scatter(h->tree.N, 3, h->perm, cp, vel, false, h->ctx->stream);
}

}  // namespace

extern "C" {

int sbslpm_irreg_upward_pass(sbslpm_irreg* h, const double* area, const double* pots, double* proxy_source_pots) {
  if (!h || !area || !pots || !proxy_source_pots) return h ? fail(h->ctx, SBSLPM_E_ARG, "irreg_upward_pass: null") : SBSLPM_E_ARG;
  irreg_upward(h, area, pots, h->q[0], proxy_source_pots);
  return launched(h->ctx);
}

int sbslpm_irreg_downward_pass_3(sbslpm_irreg* h, double* proxy_1, double* proxy_2, double* proxy_3, double* sols_1,
                                 double* sols_2, double* sols_3) {
  if (!h || !proxy_1 || !proxy_2 || !proxy_3 || !sols_1 || !sols_2 || !sols_3)
    return h ? fail(h->ctx, SBSLPM_E_ARG, "irreg_downward_pass_3: null") : SBSLPM_E_ARG;
  double* px[3] = {proxy_1, proxy_2, proxy_3};
  irreg_downward(h, px, nullptr, h->part);
  const double* cp[3] = {h->part[0], h->part[1], h->part[2]};
  double* so[3] = {sols_1, sols_2, sols_3};
  launch_scatter(h->tree.N, 3, h->perm, cp, so, true, h->ctx->stream);
  return launched(h->ctx);
}

int sbslpm_irreg_swe_vel_interactions(sbslpm_irreg* h, double* target_pots_1, double* target_pots_2,
                                      double* target_pots_3, double* vel_x, double* vel_y, double* vel_z,
                                      const double* proxy_vors, const double* proxy_divs, const double* vors,
                                      const double* divs, const double* area) {
  if (!h || !target_pots_1 || !target_pots_2 || !target_pots_3 || !vel_x || !vel_y || !vel_z || !proxy_vors ||
      !proxy_divs || !vors || !divs || !area)
    return h ? fail(h->ctx, SBSLPM_E_ARG, "irreg_swe_vel_interactions: null") : SBSLPM_E_ARG;
  const IrregTree& t = h->tree;
  hipStream_t s = h->ctx->stream;
  launch_gather_charges(t.N, h->perm, area, vors, h->q[0], s);
  launch_gather_charges(t.N, h->perm, area, divs, h->q[1], s);
  irreg_swe_pairs(h, h->q[0], proxy_vors, h->q[1], proxy_divs);
  double* tp[3] = {target_pots_1, target_pots_2, target_pots_3};
  const double* cx[3] = {h->prox[0], h->prox[1], h->prox[2]};
  launch_add((int64_t)t.P * t.n, 3, tp, cx, s);
  const double* cp[3] = {h->part[0], h->part[1], h->part[2]};
  double* vo[3] = {vel_x, vel_y, vel_z};
  launch_scatter(t.N, 3, h->perm, cp, vo, true, s);
  return launched(h->ctx);
}

int sbslpm_irreg_swe_velocity(sbslpm_irreg* h, const double* area, const double* vors, const double* divs,
                              double* vel_x, double* vel_y, double* vel_z) {
  if (!h || !area || !vors || !divs || !vel_x || !vel_y || !vel_z)
    return h ? fail(h->ctx, SBSLPM_E_ARG, "irreg_swe_velocity: null") : SBSLPM_E_ARG;
  double* vo[3] = {vel_x, vel_y, vel_z};
  irreg_swe_vel(h, area, vors, divs, vo);
  return launched(h->ctx);
}

int sbslpm_irreg_bve_velocity(sbslpm_irreg* h, const double* area, const double* vor, double* vel_x, double* vel_y,
                              double* vel_z) {
  if (!h || !area || !vor || !vel_x || !vel_y || !vel_z)
    return h ? fail(h->ctx, SBSLPM_E_ARG, "irreg_bve_velocity: null") : SBSLPM_E_ARG;
  const IrregTree& t = h->tree;
  hipStream_t s = h->ctx->stream;
  irreg_upward(h, area, vor, h->q[0], h->qx[0]);
  launch_source_records64(h->dt, h->q[0], h->qx[0], nullptr, nullptr, true, h->rec64, s);
  launch_zero_rows(h->zero_rows, h->nzero, t.n, 3, h->prox, s);
  launch_bve_interact64(h->dt, h->tiles, h->ntiles, h->sets, h->rec64, h->ctx->prm.kernel_eps, h->part, h->prox, s,
                        true);
  const double* cp[3] = {h->part[0], h->part[1], h->part[2]};
  irreg_downward(h, h->prox, cp, h->part, true);
  double* vo[3] = {vel_x, vel_y, vel_z};
  launch_scatter(t.N, 3, h->perm, cp, vo, false, s);
  return launched(h->ctx);
}

int sbslpm_interp_points(sbslpm_ctx* c, int64_t n, const double* x, const double* y, const double* z, int nfields,
                         const double* const* fields, double* const* out) {
  if (!c || n < 0 || (n > 0 && (!x || !y || !z)) || nfields < 1 || nfields > max_interp_fields() || !fields || !out)
    return fail(c, SBSLPM_E_ARG, "interp_points: bad argument");
  for (int q = 0; q < nfields; ++q)
    if (!fields[q] || !out[q]) return fail(c, SBSLPM_E_ARG, "interp_points: null field");
  launch_interp_points(c->dt, n, x, y, z, nfields, fields, out, c->stream);
  return launched(c);
}

int sbslpm_interp_points_adjoint(sbslpm_ctx* c, int64_t n, const double* x, const double* y, const double* z,
                                 int nfields, const double* const* out_bar, double* const* field_bar) {
  if (!c || n < 0 || (n > 0 && (!x || !y || !z)) || nfields < 1 || nfields > max_interp_fields() || !out_bar ||
      !field_bar)
    return fail(c, SBSLPM_E_ARG, "interp_points_adjoint: bad argument");
  for (int q = 0; q < nfields; ++q) {
    if (!out_bar[q] || !field_bar[q]) return fail(c, SBSLPM_E_ARG, "interp_points_adjoint: null field");
    launch_interp_points_adjoint(c->dt, n, x, y, z, out_bar[q], field_bar[q], c->stream);
  }
  return launched(c);
}

int sbslpm_xyz_vel_to_uv_vel(sbslpm_ctx* c, int64_t n, const double* x, const double* y, const double* z,
                             const double* vx, const double* vy, const double* vz, double* u, double* v) {
  if (!c || n < 0 || (n > 0 && (!x || !y || !z || !vx || !vy || !vz || !u || !v)))
    return fail(c, SBSLPM_E_ARG, "xyz_vel_to_uv_vel: bad argument");
  launch_xyz_vel_to_uv(n, x, y, z, vx, vy, vz, u, v, c->stream);
  return launched(c);
}

}  // extern "C"

namespace {

// A lat-lon grid the operators accept: at least 3 x 3 points and a positive spacing.
int ll_grid_check(sbslpm_ctx* c, int nlat, int nlon, double dx, const char* what) {
  if (nlat < 3 || nlon < 3 || !(dx > 0.0) || !std::isfinite(dx))
    return fail(c, SBSLPM_E_ARG,
                std::string(what) + ": lat_count and lon_count must be >= 3 and grid_spacing positive");
  return SBSLPM_OK;
}

// Null and aliasing checks: every pointer set, no output equal to an input or to
// another output (the stencils read neighbouring points while writing).
int ll_ptr_check(sbslpm_ctx* c, const char* what, std::initializer_list<const void*> in,
                 std::initializer_list<const void*> out) {
  for (const void* p : in)
    if (!p) return fail(c, SBSLPM_E_ARG, std::string(what) + ": null array");
  for (auto o = out.begin(); o != out.end(); ++o) {
    if (!*o) return fail(c, SBSLPM_E_ARG, std::string(what) + ": null array");
    for (const void* p : in)
      if (p == *o) return fail(c, SBSLPM_E_ARG, std::string(what) + ": an output aliases an input");
    for (auto q = out.begin(); q != o; ++q)
      if (*q == *o) return fail(c, SBSLPM_E_ARG, std::string(what) + ": two outputs alias");
  }
  return SBSLPM_OK;
}

int irreg_ll_check(sbslpm_irreg* h, int nlat, int nlon, double dx, const char* what) {
  if (int rc = ll_grid_check(h->ctx, nlat, nlon, dx, what)) return rc;
  if ((int64_t)nlat * nlon != h->tree.N)
    return fail(h->ctx, SBSLPM_E_ARG,
                std::string(what) + ": the grid has " + std::to_string((int64_t)nlat * nlon) +
                    " points, the tree " + std::to_string(h->tree.N));
  return SBSLPM_OK;
}

}  // namespace

extern "C" {

int sbslpm_ll_tendencies(sbslpm_ctx* c, int nlat, int nlon, double dx, double omega, const double* x,
                         const double* y, const double* z, const double* vors, const double* divs,
                         const double* height, const double* topo, const double* vx, const double* vy,
                         const double* vz, double* vor_t, double* div_t, double* h_t) {
  if (!c) return SBSLPM_E_ARG;
  if (int rc = ll_grid_check(c, nlat, nlon, dx, "ll_tendencies")) return rc;
  if (int rc = ll_ptr_check(c, "ll_tendencies", {x, y, z, vors, divs, height, topo, vx, vy, vz}, {vor_t, div_t, h_t}))
    return rc;
  const double* vel[3] = {vx, vy, vz};
  launch_ll_tendencies(ll_grid(nlat, nlon, dx), omega, x, y, z, vors, divs, height, topo, vel, vor_t, div_t, h_t,
                       c->stream);
  return launched(c);
}

int sbslpm_ll_polar_filter(sbslpm_ctx* c, int nlat, int nlon, double dx, const double* x, const double* y,
                           const double* z, double* field) {
  if (!c) return SBSLPM_E_ARG;
  if (int rc = ll_grid_check(c, nlat, nlon, dx, "ll_polar_filter")) return rc;
  if (int rc = ll_ptr_check(c, "ll_polar_filter", {x, y, z}, {field})) return rc;
  launch_ll_polar_filter(ll_grid(nlat, nlon, dx), x, y, z, field, c->stream);
  return launched(c);
}

int sbslpm_ll_arrival_values(sbslpm_ctx* c, int nlat, int nlon, double dx, double dt, const double* x,
                             const double* y, const double* z, const double* vors, const double* divs,
                             const double* height, const double* vx, const double* vy, const double* vz,
                             const double* vor_t, const double* div_t, double* a_vor, double* a_div,
                             double* a_h) {
  if (!c) return SBSLPM_E_ARG;
  if (int rc = ll_grid_check(c, nlat, nlon, dx, "ll_arrival_values")) return rc;
  if (int rc = ll_ptr_check(c, "ll_arrival_values", {x, y, z, vors, divs, height, vx, vy, vz, vor_t, div_t},
                            {a_vor, a_div, a_h}))
    return rc;
  if (int rc = ensure_scratch(c, 4)) return rc;  // the four pole means
  const LlGrid g = ll_grid(nlat, nlon, dx);
  const double* vel[3] = {vx, vy, vz};
  launch_ll_pole_means(g, dt, vors, divs, height, vor_t, c->scratch, c->stream);
  launch_ll_arrival(g, dt, x, y, z, vors, divs, height, vel, vor_t, div_t, c->scratch, a_vor, a_div, a_h, c->stream);
  return launched(c);
}

int sbslpm_ll_field_averaging(sbslpm_ctx* c, int nlat, int nlon, const double* x, const double* y, const double* z,
                              const double* field, double* new_field, double epsilon) {
  if (!c) return SBSLPM_E_ARG;
  if (int rc = ll_grid_check(c, nlat, nlon, 1.0, "ll_field_averaging")) return rc;
  if (int rc = ll_ptr_check(c, "ll_field_averaging", {x, y, z, field}, {new_field})) return rc;
  launch_ll_average(ll_grid(nlat, nlon, 1.0), x, y, z, field, new_field, epsilon, c->stream);
  return launched(c);
}

int sbslpm_ll_laplacian_diffusion(sbslpm_ctx* c, int nlat, int nlon, double dx, double dt, const double* x,
                                  const double* y, const double* z, const double* field, double* new_field) {
  if (!c) return SBSLPM_E_ARG;
  if (int rc = ll_grid_check(c, nlat, nlon, dx, "ll_laplacian_diffusion")) return rc;
  if (int rc = ll_ptr_check(c, "ll_laplacian_diffusion", {x, y, z, field}, {new_field})) return rc;
  launch_ll_diffuse(ll_grid(nlat, nlon, dx), dt, x, y, z, field, new_field, c->stream);
  return launched(c);
}

int sbslpm_irreg_swe_vel_tend_ll(sbslpm_irreg* h, int nlat, int nlon, double dx, double omega, const double* x,
                                 const double* y, const double* z, const double* area, const double* vors,
                                 const double* divs, const double* height, const double* topo, double* vx,
                                 double* vy, double* vz, double* vor_t, double* div_t, double* h_t) {
  if (!h) return SBSLPM_E_ARG;
  sbslpm_ctx* c = h->ctx;
  if (int rc = irreg_ll_check(h, nlat, nlon, dx, "irreg_swe_vel_tend_ll")) return rc;
  if (int rc = ll_ptr_check(c, "irreg_swe_vel_tend_ll", {x, y, z, area, vors, divs, height, topo},
                            {vx, vy, vz, vor_t, div_t, h_t}))
    return rc;
  const LlGrid g = ll_grid(nlat, nlon, dx);
  double* vel[3] = {vx, vy, vz};
  irreg_swe_vel(h, area, vors, divs, vel);
  const double* cv[3] = {vx, vy, vz};
  launch_ll_tendencies(g, omega, x, y, z, vors, divs, height, topo, cv, vor_t, div_t, h_t, c->stream);
  launch_ll_polar_filter(g, x, y, z, h_t, c->stream);
  return launched(c);
}

int sbslpm_irreg_swe_back_rk4_step_ll(sbslpm_irreg* h, int nlat, int nlon, double dx, double dt, double omega,
                                      const double* x, const double* y, const double* z, const double* area,
                                      double* vors, double* divs, double* height, const double* topo, double* vx,
                                      double* vy, double* vz, double time) {
  (void)time;
  if (!h) return SBSLPM_E_ARG;
  sbslpm_ctx* c = h->ctx;
  if (int rc = irreg_ll_check(h, nlat, nlon, dx, "irreg_swe_back_rk4_step_ll")) return rc;
  if (int rc = ll_ptr_check(c, "irreg_swe_back_rk4_step_ll", {x, y, z, area, topo}, {vors, divs, height, vx, vy, vz}))
    return rc;
  const int64_t n = h->tree.N;
  if (!h->ll_ws) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (c->stream && hipStreamIsCapturing(c->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
      return fail(c, SBSLPM_E_ARG, "irreg_swe_back_rk4_step_ll: call once outside stream capture first (workspace)");
    HIP_TRY(c, hipMalloc(&h->ll_ws, (5 * n + 8) * sizeof(double)));
  }
  double* vor_t = h->ll_ws;
  double* div_t = vor_t + n;
  double* a_vor = div_t + n;
  double* a_div = a_vor + n;
  double* a_h = a_div + n;
  double* pole = a_h + n;
  const LlGrid g = ll_grid(nlat, nlon, dx);
  double* vel[3] = {vx, vy, vz};
  irreg_swe_vel(h, area, vors, divs, vel);
  const double* cv[3] = {vx, vy, vz};
  launch_ll_tendencies(g, omega, x, y, z, vors, divs, height, topo, cv, vor_t, div_t, nullptr, c->stream);
  launch_ll_pole_means(g, dt, vors, divs, height, vor_t, pole, c->stream);
  launch_ll_arrival(g, dt, x, y, z, vors, divs, height, cv, vor_t, div_t, pole, a_vor, a_div, a_h, c->stream);
  launch_ll_smooth(g, dt, x, y, z, a_vor, a_div, a_h, 0.16, vors, divs, height, c->stream);
  launch_ll_polar_filter(g, x, y, z, divs, c->stream);
  return launched(c);
}

int sbslpm_bve_back_rk4_step(sbslpm_ctx* c, const double* x, const double* y, const double* z, const double* area,
                             double* vor, double* proxy_source_vors, double* vx, double* vy, double* vz,
                             double* tracers, double time) {
  (void)time;
  (void)proxy_source_vors;  // the context owns its proxy charge workspace
  if (!c || !x || !y || !z || !area || !vor || !vx || !vy || !vz || (c->prm.tracer_count > 0 && !tracers))
    return fail(c, SBSLPM_E_ARG, "bve_back_rk4_step: bad argument");
  double* vel0[3] = {vx, vy, vz};
  // Graph capture needs a non-default stream; on the default stream the step runs eagerly.
  if (!(c->prm.flags & SBSLPM_FLAG_GRAPH) || !c->stream) return bve_step_body(c, c->w1, 1, x, y, z, area, vor, vel0, tracers);

  const void* key[10] = {x, y, z, area, vor, vx, vy, vz, tracers, (const void*)(intptr_t)c->prm.precision};
  if (!c->graph || std::memcmp(key, c->graph_key, sizeof(key)) != 0) {
    if (c->graph) {
      hipGraphExecDestroy(c->graph);
      c->graph = nullptr;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    int rc = bve_step_body(c, c->w1, 1, x, y, z, area, vor, vel0, tracers);
    c->capturing = false;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(c, SBSLPM_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
      c->graph = nullptr;
      return fail(c, SBSLPM_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    std::memcpy(c->graph_key, key, sizeof(key));
  }
  HIP_TRY(c, hipGraphLaunch(c->graph, c->stream));
  return SBSLPM_OK;
}

// Adjoint workspace and the source-major interaction list (built on the host
// on the first adjoint call).
static int ensure_adjoint(sbslpm_ctx* c) {
  if (c->adj_arena) return SBSLPM_OK;
  const Tree& t = c->tree;
  SourceMajor sm;
  build_source_major(t, sm);
  const size_t NP = (size_t)t.npart, Pn = (size_t)t.P * t.n;
  Arena A;
  const size_t o_tiles = A.take<TargetTile>(sm.tiles.size()), o_tg = A.take<SourceSet>(sm.targets.size()),
               o_zr = A.take<int>(sm.zero_rows.size()), o1 = A.take<double>(NP), o_pp = A.take<double>(3 * NP),
               o_px = A.take<double>(3 * Pn), o_tr = A.take<char>(t.U * adj_target_record_bytes()),
               o_sr = A.take<SrcRec64>(t.U), o_qbp = A.take<double>(NP), o_qbx = A.take<double>(Pn),
               o_qb = A.take<double>(NP), o_dbp = A.take<double>(3 * NP), o_dbx = A.take<double>(3 * Pn),
               o_v = A.take<double>(12 * NP), o_vb = A.take<double>(12 * NP), total = A.take<char>(256);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  char* b = nullptr;
  if (hipMalloc(&b, total) != hipSuccess) return fail(c, SBSLPM_E_NOMEM, "adjoint workspace: out of device memory");
  c->adj_arena = b;
  c->adj_tiles = (TargetTile*)(b + o_tiles);
  c->adj_targets = (SourceSet*)(b + o_tg);
  c->adj_zero_rows = (int*)(b + o_zr);
  c->adj_ntiles = (int)sm.tiles.size();
  c->adj_nzero = (int)sm.zero_rows.size();
  c->adj_ones = (double*)(b + o1);
  c->adj_trec = b + o_tr;
  c->adj_srec = (SrcRec64*)(b + o_sr);
  c->adj_qbp = (double*)(b + o_qbp);
  c->adj_qbx = (double*)(b + o_qbx);
  c->adj_qb = (double*)(b + o_qb);
  for (int f = 0; f < 3; ++f) {
    c->adj_psip[f] = (double*)(b + o_pp) + f * NP;
    c->adj_psix[f] = (double*)(b + o_px) + f * Pn;
    c->adj_dbp[f] = (double*)(b + o_dbp) + f * NP;
    c->adj_dbx[f] = (double*)(b + o_dbx) + f * Pn;
  }
  for (int i = 0; i < 12; ++i) {
    c->adj_v[i] = (double*)(b + o_v) + i * NP;
    c->adj_vb[i] = (double*)(b + o_vb) + i * NP;
  }
  int rc = upload(c, c->adj_tiles, sm.tiles.data(), sm.tiles.size() * sizeof(TargetTile));
  if (!rc) rc = upload(c, c->adj_targets, sm.targets.data(), sm.targets.size() * sizeof(SourceSet));
  if (!rc) rc = upload(c, c->adj_zero_rows, sm.zero_rows.data(), sm.zero_rows.size() * sizeof(int));
  if (rc) return rc;
  launch_fill_value(c->adj_ones, (int64_t)NP, 1.0, c->stream);
  return launched(c);
}

// Exact transpose of velocity() in FP64 for the particle-velocity cotangent
// vbar: assigns adj_qb = dL/dq_part (q_part = area * vor_eff) and, when the
// sources are displaced (dpart, dprox given), adj_dbp = dL/d disp_part
// including the path through the interpolated proxy displacements.  The
// charges (area, vor_eff) enter only the displacement part.
static int fmm_adjoint(sbslpm_ctx* c, const double* area, const double* vor_eff, const double* const* dpart,
                const double* const* dprox, const double* const* vbar) {
  const Tree& t = c->tree;
  const bool pos = dpart != nullptr;
  {  // the cotangent's anterpolation and (displaced sources) the charge's upward pass share their launches
    UpCharge up[4];
    anterpolation_sets(c, vbar, 3, up);
    up[3] = {area, vor_eff, c->qpart, c->qproxy};
    launch_upward_sets(c->dt, up, pos ? 4 : 3, c->stream);
  }
  const double* psix[3] = {c->adj_psix[0], c->adj_psix[1], c->adj_psix[2]};
  const bool fp32 = c->prm.precision != SBSLPM_FP64;  // mixed contexts take the FP32 pair sums
  if (fp32)
    launch_adj_target_records32(c->dt, vbar, psix, c->adj_trec, c->stream);
  else
    launch_adj_target_records(c->dt, vbar, psix, c->adj_trec, c->stream);
  // source positions (and, for the displacement part, charges) as the forward records
  launch_source_records64(c->dt, c->qpart, c->qproxy, dpart, dprox, true, c->adj_srec, c->stream);
  launch_zero_rows(c->adj_zero_rows, c->adj_nzero, t.n, 1, &c->adj_qbx, c->stream);
  if (pos) launch_zero_rows(c->adj_zero_rows, c->adj_nzero, t.n, 3, c->adj_dbx, c->stream);
  (fp32 ? launch_bve_adjoint32 : launch_bve_adjoint64)(c->dt, c->adj_tiles, c->adj_ntiles, c->adj_targets,
                                                       c->adj_trec, c->adj_srec, c->prm.kernel_eps, dpart, dprox,
                                                       c->adj_qbp, c->adj_qbx, c->adj_dbp, c->adj_dbx, c->stream,
                                                       c->shape);
  // charges: particle sources directly, proxy sources through P2M/M2M, whose
  // transposes are L2L/L2P
  const double* qbp[1] = {c->adj_qbp};
  launch_downward_pass(c->dt, 1, &c->adj_qbx, qbp, &c->adj_qb, c->stream, true);
  if (pos) {
    const double* dbx[3] = {c->adj_dbx[0], c->adj_dbx[1], c->adj_dbx[2]};
    launch_disp_interp_adjoint(c->dt, dbx, c->adj_dbp, c->stream);
  }
  return launched(c);
}

int sbslpm_bve_velocity_adjoint(sbslpm_ctx* c, const double* area, const double* ubx, const double* uby,
                                const double* ubz, double* vor_bar) {
  if (!c || !area || !ubx || !uby || !ubz || !vor_bar) return fail(c, SBSLPM_E_ARG, "bve_velocity_adjoint: null pointer");
  int rc = ensure_adjoint(c);
  if (rc) return rc;
  const double* ub[3] = {ubx, uby, ubz};
  rc = fmm_adjoint(c, area, nullptr, nullptr, nullptr, ub);
  if (rc) return rc;
  launch_fill_zero(vor_bar, c->tree.npart, c->stream);
  launch_stage_adjoint(c->dt, nullptr, nullptr, nullptr, nullptr, 0.0, 0.0, area, c->adj_qb, nullptr, vor_bar,
                       nullptr, c->stream);
  return launched(c);
}

}  // extern "C"

namespace {

// The four stage velocities of one bve_back_rk4_step into V[0..11] (v0..v3),
// with the context's precision and kernels.
int step_stages(sbslpm_ctx* c, const double* x, const double* y, const double* z, const double* area, const double* vor,
                double* const* V) {
  const Work& w = c->w1;
  const double dt = c->prm.delta_t, om = c->prm.omega;
  const double taus[3] = {0.5 * dt, 0.5 * dt, dt};
  const double* cdp[3] = {w.disppart[0], w.disppart[1], w.disppart[2]};
  const double* cdx[3] = {w.dispproxy[0], w.dispproxy[1], w.dispproxy[2]};
  int rc = velocity(c, w, 1, area, vor, nullptr, nullptr, V);
  for (int s = 0; s < 3 && rc == SBSLPM_OK; ++s) {
    const double* vin[3] = {V[3 * s], V[3 * s + 1], V[3 * s + 2]};
    launch_bve_displacement(c->dt, x, y, z, vin, vor, taus[s], om, w.newvor, w.disppart, w.dispproxy, c->stream);
    rc = velocity(c, w, 1, area, w.newvor, cdp, cdx, V + 3 * (s + 1));
  }
  return rc;
}

// Reverse pass of one step given its stage velocities V (see
// sbslpm_bve_back_rk4_step_adjoint).
int step_reverse(sbslpm_ctx* c, const double* x, const double* y, const double* z, const double* area,
                 const double* vor, const double* const* V, const double* new_vor_bar, double* vor_bar) {
  const Tree& t = c->tree;
  const Work& w = c->w1;
  const double dt = c->prm.delta_t, om = c->prm.omega;
  const double taus[3] = {0.5 * dt, 0.5 * dt, dt};
  double* const* VB = c->adj_vb;
  const double* cdp[3] = {w.disppart[0], w.disppart[1], w.disppart[2]};
  const double* cdx[3] = {w.dispproxy[0], w.dispproxy[1], w.dispproxy[2]};
  int rc = SBSLPM_OK;
  // remap and departure search
  {  // vor_bar and the v0..v2 cotangents (contiguous) in one launch
    double* zero[2] = {vor_bar, VB[0]};
    const int64_t count[2] = {t.npart, 9 * t.npart};
    launch_fill_zero_sets(zero, count, 2, c->stream);
  }
  launch_departure_adjoint(c->dt, x, y, z, V, vor, new_vor_bar, dt, c->prm.rk4_combine, om, VB, vor_bar, c->stream);
  // stages 3, 2, 1: velocity at displaced sources, then the displacement update
  for (int s = 2; s >= 0 && rc == SBSLPM_OK; --s) {
    const double* vin[3] = {V[3 * s], V[3 * s + 1], V[3 * s + 2]};
    launch_bve_displacement(c->dt, x, y, z, vin, vor, taus[s], om, w.newvor, w.disppart, w.dispproxy, c->stream);
    const double* vb[3] = {VB[3 * (s + 1)], VB[3 * (s + 1) + 1], VB[3 * (s + 1) + 2]};
    rc = fmm_adjoint(c, area, w.newvor, cdp, cdx, vb);
    const double* db[3] = {c->adj_dbp[0], c->adj_dbp[1], c->adj_dbp[2]};
    launch_stage_adjoint(c->dt, x, y, z, vin, taus[s], om, area, c->adj_qb, db, vor_bar, VB + 3 * s, c->stream);
  }
  if (rc) return rc;
  // stage 0: undisplaced sources
  const double* vb0[3] = {VB[0], VB[1], VB[2]};
  rc = fmm_adjoint(c, area, nullptr, nullptr, nullptr, vb0);
  if (rc) return rc;
  launch_stage_adjoint(c->dt, nullptr, nullptr, nullptr, nullptr, 0.0, 0.0, area, c->adj_qb, nullptr, vor_bar, nullptr,
                       c->stream);
  return launched(c);
}

}  // namespace

extern "C" {

int sbslpm_bve_back_rk4_step_adjoint(sbslpm_ctx* c, const double* x, const double* y, const double* z,
                                     const double* area, const double* vor, const double* new_vor_bar,
                                     double* vor_bar) {
  if (!c || !x || !y || !z || !area || !vor || !new_vor_bar || !vor_bar)
    return fail(c, SBSLPM_E_ARG, "bve_back_rk4_step_adjoint: null pointer");
  if (vor_bar == vor || vor_bar == new_vor_bar)
    return fail(c, SBSLPM_E_ARG, "bve_back_rk4_step_adjoint: vor_bar must not alias the inputs");
  int rc = ensure_adjoint(c);
  if (!rc) rc = step_stages(c, x, y, z, area, vor, c->adj_v);
  if (rc) return rc;
  const double* cv[12];
  for (int i = 0; i < 12; ++i) cv[i] = c->adj_v[i];
  return step_reverse(c, x, y, z, area, vor, cv, new_vor_bar, vor_bar);
}

int sbslpm_bve_back_rk4_step_record(sbslpm_ctx* c, const double* x, const double* y, const double* z,
                                    const double* area, const double* vor, double* vor_out, double* stages) {
  if (!c || !x || !y || !z || !area || !vor || !vor_out || !stages)
    return fail(c, SBSLPM_E_ARG, "bve_back_rk4_step_record: null pointer");
  if (vor_out == vor) return fail(c, SBSLPM_E_ARG, "bve_back_rk4_step_record: vor_out must not alias vor");
  const int64_t N = c->tree.npart;
  double* V[12];
  for (int i = 0; i < 12; ++i) V[i] = stages + i * N;
  int rc = step_stages(c, x, y, z, area, vor, V);
  if (rc) return rc;
  const Work& w = c->w1;
  const double* cv[12];
  for (int i = 0; i < 12; ++i) cv[i] = V[i];
  launch_bve_departure(c->dt, x, y, z, cv, c->prm.delta_t, c->prm.rk4_combine, w.dep[0], w.dep[1], w.dep[2], c->stream);
  launch_bve_departure_to_target(c->dt, z, w.dep[0], w.dep[1], w.dep[2], vor, nullptr, 0, c->prm.omega, vor_out, nullptr,
                                 c->stream);
  return launched(c);
}

int sbslpm_bve_back_rk4_step_adjoint_stages(sbslpm_ctx* c, const double* x, const double* y, const double* z,
                                            const double* area, const double* vor, const double* stages,
                                            const double* new_vor_bar, double* vor_bar) {
  if (!c || !x || !y || !z || !area || !vor || !stages || !new_vor_bar || !vor_bar)
    return fail(c, SBSLPM_E_ARG, "bve_back_rk4_step_adjoint_stages: null pointer");
  if (vor_bar == vor || vor_bar == new_vor_bar)
    return fail(c, SBSLPM_E_ARG, "bve_back_rk4_step_adjoint_stages: vor_bar must not alias the inputs");
  int rc = ensure_adjoint(c);
  if (rc) return rc;
  const int64_t N = c->tree.npart;
  const double* cv[12];
  for (int i = 0; i < 12; ++i) cv[i] = stages + i * N;
  return step_reverse(c, x, y, z, area, vor, cv, new_vor_bar, vor_bar);
}

int sbslpm_bve_departure_to_target_adjoint(sbslpm_ctx* c, const double* dep_x, const double* dep_y,
                                           const double* dep_z, const double* new_vor_bar, double* vor_bar) {
  if (!c || !dep_x || !dep_y || !dep_z || !new_vor_bar || !vor_bar)
    return fail(c, SBSLPM_E_ARG, "bve_departure_to_target_adjoint: null pointer");
  launch_remap_adjoint(c->dt, dep_x, dep_y, dep_z, new_vor_bar, vor_bar, c->stream);
  return launched(c);
}

// Ensemble workspace for nmem members (grown on demand, never shrunk).
static int ensure_ensemble(sbslpm_ctx* c, int nmem) {
  const bool mrhs = c->prm.precision == SBSLPM_MIXED && (c->prm.flags & SBSLPM_FLAG_MULTI_RHS) && nmem > 1;
  // reuse only a workspace whose record buffers match the context's current
  // precision (set_precision may have changed it since) and that has the
  // multi-RHS columns when they are needed
  if (c->ens_arena && c->wens.members >= nmem && c->wens.precision == c->prm.precision && (c->wens.mrhs || !mrhs))
    return SBSLPM_OK;
  if (c->capturing) return fail(c, SBSLPM_E_ARG, "ensemble workspace: cannot allocate during graph capture");
  const Tree& t = c->tree;
  const size_t NP = (size_t)t.npart * nmem, Pn = (size_t)t.P * t.n * nmem, U = (size_t)t.U * nmem,
               U16 = (size_t)t.U16 * nmem;
  Arena A;
  const size_t o_r64 = A.take<SrcRec64>(c->prm.precision == SBSLPM_FP64 ? U : 0);
  const size_t o_r32 = A.take<SrcRec32>(c->prm.precision == SBSLPM_MIXED_VALU ? U : 0);
  const size_t o_rmf = A.take<SrcRecMfma>(c->prm.precision == SBSLPM_MIXED ? U16 : 0);
  const size_t o_qc = A.take<float>(mrhs ? (size_t)t.U16 * 16 * mrhs_chunks(nmem, 2) : 0);
  const size_t o_qp = A.take<double>(NP), o_qx = A.take<double>(Pn);
  const size_t o_pp = A.take<double>(3 * NP), o_px = A.take<double>(3 * Pn);
  const size_t o_dp = A.take<double>(3 * NP), o_dx = A.take<double>(3 * Pn);
  const size_t o_vs = A.take<double>(9 * NP), o_nv = A.take<double>(NP), o_dep = A.take<double>(3 * NP);
  const size_t o_vn = A.take<double>(NP), total = A.take<char>(256);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (c->ens_graph) {
    hipGraphExecDestroy(c->ens_graph);
    c->ens_graph = nullptr;
  }
  if (c->ens_arena) hipFree(c->ens_arena);
  c->ens_arena = nullptr;
  c->wens = Work{};
  char* b = nullptr;
  if (hipMalloc(&b, total) != hipSuccess) return fail(c, SBSLPM_E_NOMEM, "ensemble workspace: out of device memory");
  c->ens_arena = b;
  Work& w = c->wens;
  w.members = nmem;
  w.precision = c->prm.precision;
  w.mrhs = mrhs;
  w.rec64 = (SrcRec64*)(b + o_r64); w.rec32 = (SrcRec32*)(b + o_r32); w.recmf = (SrcRecMfma*)(b + o_rmf);
  w.qcol = mrhs ? (float*)(b + o_qc) : nullptr;
  w.qpart = (double*)(b + o_qp); w.qproxy = (double*)(b + o_qx);
  for (int f = 0; f < 3; ++f) {
    w.phipart[f] = (double*)(b + o_pp) + f * NP;
    w.phiproxy[f] = (double*)(b + o_px) + f * Pn;
    w.disppart[f] = (double*)(b + o_dp) + f * NP;
    w.dispproxy[f] = (double*)(b + o_dx) + f * Pn;
    w.dep[f] = (double*)(b + o_dep) + f * NP;
  }
  for (int k = 0; k < 9; ++k) w.vstage[k] = (double*)(b + o_vs) + k * NP;
  w.newvor = (double*)(b + o_nv);
  w.vor_next = (double*)(b + o_vn);
  return SBSLPM_OK;
}

// Stacked per-member arrays are [member][...] with the member as the slowest
// index; component arrays of the workspace are [component][member][...].  The
// kernels index member j of a component array at j * (member size), so the
// workspace keeps every component's members contiguous: phipart[f] + j * NP.
int sbslpm_bve_back_rk4_step_ensemble(sbslpm_ctx* c, int nmem, const double* x, const double* y, const double* z,
                                      const double* area, double* vor, double* vx, double* vy, double* vz,
                                      double time) {
  (void)time;
  if (!c || nmem < 1 || nmem > 65535 || !x || !y || !z || !area || !vor || !vx || !vy || !vz)
    return fail(c, SBSLPM_E_ARG, "bve_back_rk4_step_ensemble: bad argument");
  if (c->tree.cfg.levels < 2) return fail(c, SBSLPM_E_ARG, "bve_back_rk4_step_ensemble: needs levels >= 2");
  int rc = ensure_ensemble(c, nmem);
  if (rc) return rc;
  double* vel0[3] = {vx, vy, vz};
  if (!(c->prm.flags & SBSLPM_FLAG_GRAPH) || !c->stream)
    return bve_step_body(c, c->wens, nmem, x, y, z, area, vor, vel0, nullptr);
  const void* key[10] = {x,  y,  z,  area, vor, vx, vy, vz, (const void*)(intptr_t)nmem,
                         (const void*)(intptr_t)c->prm.precision};
  if (!c->ens_graph || std::memcmp(key, c->ens_key, sizeof(key)) != 0) {
    if (c->ens_graph) {
      hipGraphExecDestroy(c->ens_graph);
      c->ens_graph = nullptr;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    rc = bve_step_body(c, c->wens, nmem, x, y, z, area, vor, vel0, nullptr);
    c->capturing = false;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(c, SBSLPM_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->ens_graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
      c->ens_graph = nullptr;
      return fail(c, SBSLPM_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    std::memcpy(c->ens_key, key, sizeof(key));
  }
  HIP_TRY(c, hipGraphLaunch(c->ens_graph, c->stream));
  return SBSLPM_OK;
}

int sbslpm_autotune_ensemble(sbslpm_ctx* c, int nmem, const double* area, const double* vor, double* report,
                             int* ncand) {
  if (!c || !area || !vor) return fail(c, SBSLPM_E_ARG, "autotune_ensemble: null argument");
  if (c->capturing) return fail(c, SBSLPM_E_ARG, "autotune_ensemble: not allowed during graph capture");
  if (c->prm.precision != SBSLPM_MIXED || !(c->prm.flags & SBSLPM_FLAG_MULTI_RHS))
    return fail(c, SBSLPM_E_ARG, "autotune_ensemble: needs a mixed-precision context with SBSLPM_FLAG_MULTI_RHS");
  if (nmem < 2 || nmem > 65535) return fail(c, SBSLPM_E_ARG, "autotune_ensemble: needs 2 <= members <= 65535");
  int rc = ensure_ensemble(c, nmem);
  if (rc) return rc;
  const Work& w = c->wens;
  launch_upward_fused(c->dt, area, vor, w.qpart, w.qproxy, c->stream, nmem);
  launch_source_records_mfma(c->dt, w.qpart, w.qproxy, nullptr, nullptr, w.recmf, c->stream, 1);
  hipEvent_t e0, e1;
  HIP_TRY(c, hipEventCreate(&e0));
  HIP_TRY(c, hipEventCreate(&e1));
  const int waves[3] = {1, 2, 4};
  PairShape best = c->shape;
  int best_cb = mrhs_chunks_per_wave(nmem);
  float best_ms = 1e30f;
  int k = 0;
  for (int cb = 1; cb <= 2; ++cb) {
    launch_source_columns_mrhs(c->dt, w.qpart, w.qproxy, w.recmf, nmem, cb, w.qcol, c->stream);
    for (int xcd = 0; xcd < 3; ++xcd)
      for (int wi = 0; wi < 3; ++wi) {
        const PairShape sh = full_list_shape(c->tree, waves[wi], xcd, true);
        auto launch = [&] {
          launch_bve_interact_mrhs(c->dt, c->tiles, c->ntiles, c->sets, w.recmf, w.qcol, nmem, cb,
                                   c->prm.kernel_eps, w.phipart, w.phiproxy, c->stream, true, sh);
        };
        launch();  // warm-up
        hipEventRecord(e0, c->stream);
        for (int r = 0; r < 3; ++r) launch();
        hipEventRecord(e1, c->stream);
        if (hipEventSynchronize(e1) != hipSuccess) break;
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        ms /= 3;
        if (report) {
          report[4 * k] = sh.waves;
          report[4 * k + 1] = sh.xcd;
          report[4 * k + 2] = cb;
          report[4 * k + 3] = ms;
        }
        ++k;
        if (ms < best_ms) {
          best_ms = ms;
          best = sh;
          best_cb = cb;
        }
      }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (ncand) *ncand = k;
  rc = launched(c);
  if (rc) return rc;
  c->mrhs_shape = best;
  c->mrhs_cb = best_cb;
  // a captured ensemble step holds the old launch shape
  if (c->ens_graph) {
    hipGraphExecDestroy(c->ens_graph);
    c->ens_graph = nullptr;
  }
  return SBSLPM_OK;
}

}  // extern "C"
