// C-ABI host runtime of the Sinkhorn-Knopp WMD engine (declarations in
// include/wmd.h).  Owns device memory, converts the reference's V x N
// word-major CSR (common.h:38-46) into the device layout once per corpus,
// launches the gfx950 kernels of wmd_kernels.hip, and captures the streaming
// variants' iteration loop in a HIP graph.
#include "../../include/wmd.h"
#include "wmd_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

void set_err(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

}  // namespace

// Error text of the file-format readers (wmd_io.cpp) goes to the same
// thread-local message.
namespace wmd_io {
void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}
}  // namespace wmd_io

namespace {

#define HIP_TRY(expr)                                                             \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      set_err("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return WMD_ERR_HIP;                                                         \
    }                                                                             \
  } while (0)

// Grow-only device buffer.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;  // elements
  hipError_t ensure(size_t n) {
    if (n <= cap && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      cap = 0;
      if (e != hipSuccess) return e;
    }
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), bytes);
    if (e == hipSuccess) cap = std::max<size_t>(n, 1);
    return e;
  }
  hipError_t upload(const T* h, size_t n, hipStream_t s) {
    hipError_t e = ensure(n);
    if (e != hipSuccess || n == 0) return e;
    return hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

int round_up4(int x) { return (x + 3) & ~3; }

// A temporary device allocation freed when it goes out of scope.
struct ScopedDev {
  void* p = nullptr;
  ScopedDev() = default;
  ScopedDev(const ScopedDev&) = delete;
  ScopedDev& operator=(const ScopedDev&) = delete;
  ~ScopedDev() {
    if (p) (void)hipFree(p);
  }
};

// One resident launch: queries bucket[off, off + cnt) against the documents
// doc_perm[d0, d1) of one length class, with one tile shape; or, with shape
// kLongLaunch, against the documents long_docs[d0, d1) beyond the tiles.
constexpr int kLongLaunch = -1;
struct ResidentLaunch {
  int shape, cls;
  size_t off;
  int64_t cnt, d0, d1;
  double work;  // the list-scheduling weight (estimated lane instructions)
};

// Per-wave instruction estimate of one iteration of a pair in shape s with
// NE document-word slots per lane: the two tile products, the B-lane
// butterfly and quotient per slot row, and the per-slot update.
// Tile rows NE of shape s on document class cls (the kernel's template
// argument; equal for consecutive classes when the shape has more document
// lanes than a class is wide).
int tile_rows(int s, int cls) {
  const wmd::TileShape& t = wmd::kShapes[s];
  const int A = 64 * t.W / t.B;
  return (wmd::kDocClassWidth * (cls + 1) + A - 1) / A;
}

double iter_cost(int s, int cls) {
  const wmd::TileShape& t = wmd::kShapes[s];
  const int A = 64 * t.W / t.B;
  const int ne = (wmd::kDocClassWidth * (cls + 1) + A - 1) / A;
  int lb = 0;
  while ((1 << lb) < t.B) ++lb;
  const int no = (t.NI + 3) / 4;
  return (double)t.W * (ne * (2.0 * t.NI + lb + 2) + 40.0 * no + 40.0);
}

}  // namespace

// Device scratch of the allocation-free entry (wmd_context_run_query_noalloc):
// a file-scope array of kScratchBytes on every GPU the library runs on, held
// until the process ends.  Host code never passes the array's name to a
// kernel: its device address, per device, comes from hipGetSymbolAddress at
// context creation (scratch_address) and every use goes through that.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Resolves the scratch array's address on `device` (which must be current)
// once; later calls return the kept pointer.
int scratch_address(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) {
    set_err("device %d beyond the %d the scratch table covers", device, kMaxDevices);
    return WMD_ERR_UNSUPPORTED;
  }
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_err("hipGetSymbolAddress(scratch) failed on device %d: %s", device,
              hipGetErrorString(e));
      return WMD_ERR_HIP;
    }
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[device];
  return WMD_OK;
}
}  // namespace

// Side streams of a context.  The resident launches of a run fan out over
// four lanes, the context stream and three side streams: the runtime maps
// consecutive streams of a process to its (by default four) hardware queues
// in turn, so the four lanes run concurrently, and one launch's tail (its
// slowest pairs) overlaps other launches.
constexpr int kSideStreams = 3;
constexpr int kLanes = kSideStreams + 1;
constexpr int kTuneStride = wmd::kMaxQ + 1;
// Shapes the autotuner times per (document class, query bucket): the default
// and the next ones in capacity order, up to twice the default's capacity.
constexpr int kTuneCandidates = 6;
constexpr float kTuneMargin = 0.97f;
constexpr int kTuneReps = 2;  // timed launches per candidate (the fastest counts)

struct wmd_context {
  int device = 0;
  unsigned char* scratch = nullptr;  // this device's scratch array (scratch_address)
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipStream_t side[kSideStreams] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kSideStreams] = {};
  uint32_t* h_status = nullptr;  // pinned
  // Tile shape per (document class, query length) and domain (0 linear,
  // 1 log): filled by wmd_context_autotune or wmd_context_set_tuning; -1 or
  // an empty table means the smallest shape that fits.
  std::vector<int32_t> tune[2];
  // Resident launch plan (its query buckets live in plan_bucket) and the
  // solver stage captured from it.  plan_epoch moves whenever corpus,
  // queries or tuning change; a plan or graph built at an older epoch, for
  // other shape/domain choices or (graph) other kernel arguments is rebuilt.
  uint64_t plan_epoch = 1, plan_built = 0, plan_id = 0;
  int plan_forced = -2;
  bool plan_log = false;
  std::vector<ResidentLaunch> plan;
  std::vector<int32_t> h_plan_flat;  // host copy of plan_bucket
  bool plan_measured = false;        // plan weights come from a finished run
  hipGraphExec_t solve_exec = nullptr;
  uint64_t solve_plan_id = 0;
  wmd::SolveArgs solve_args{};
  // Streaming variants: the last captured chunk of iterations, replayed while
  // its kernel arguments, length and form stay the same.
  hipGraphExec_t stream_exec = nullptr;
  wmd::SolveArgs stream_args{};
  int stream_len = -1, stream_chunk = -1;
  bool stream_fused = false, stream_tol = false;

  // corpus
  bool has_corpus = false;
  int64_t V = 0, D = 0, N = 0, nnz = 0, Vt = 0;
  int nmax = 0;
  DevBuf<float> E;
  DevBuf<int32_t> touched, doc_ptr, doc_t;
  DevBuf<int32_t> doc_t1;               // doc_t + 1 (the resident solver's biased word ids)
  DevBuf<int32_t> doc_perm;             // documents sorted by length class
  DevBuf<int32_t> long_docs;            // documents longer than the resident tiles
  std::vector<int32_t> h_long;          // host copy
  int n_long = 0;
  DevBuf<uint16_t> Es;                  // matrix-core stage: BF16 planes of touched rows
  DevBuf<double> nt;                    // FP64 squared norms of touched rows
  bool split_ready = false;
  DevBuf<uint16_t> Qs;                  // BF16 planes of query words
  DevBuf<double> nqw;                   // FP64 squared norms of query words
  std::vector<int64_t> h_class_off;     // kNumDocClasses + 1 offsets into doc_perm
  std::vector<int32_t> h_doc_perm;      // host copy of doc_perm
  DevBuf<float> doc_val;

  // queries
  bool has_queries = false;
  int nq = 0;
  int qmax = 0;
  std::vector<int32_t> h_q_ptr, h_k_stride;
  std::vector<int64_t> h_k_off, h_u_off;
  int64_t k_total = 0, u_total = 0;
  DevBuf<int32_t> q_ptr, q_words, k_stride, bucket, plan_bucket;
  DevBuf<float> q_r;
  DevBuf<int64_t> k_off, u_off;
  DevBuf<wmd::StreamQuery> qmeta;
  DevBuf<int32_t> act;  // streaming: pairs still iterating (status scan's list)
  DevBuf<float> K, mrow;
  // shared distance stage: the batch's distinct query words (ascending), the
  // distinct index of every query word, pseudo-queries of 64 distinct words
  // (row stride Sdp into Dd) and the Vt x Sdp raw distances
  int Sd = 0, Sdp = 0, nqd = 0;
  DevBuf<int32_t> d_words, slot, dq_ptr, dk_stride;
  DevBuf<int64_t> dk_off;
  DevBuf<float> Dd;
  DevBuf<double> wmd;
  DevBuf<int32_t> iters;
  DevBuf<float> resid;
  // streaming state
  DevBuf<float> u, ow;
  DevBuf<uint8_t> done;
  DevBuf<uint32_t> status;
  bool ran = false;

  // learned warm start
  int n_cu = 0;
  std::vector<int32_t> h_doc_ptr, h_doc_t, h_touched;  // host copies for training batches
  DevBuf<int32_t> doc_of;               // nnz: document of each nonzero
  DevBuf<float> mlp_p, mlp_m, mlp_v;    // parameters, Adam moments
  int adam_step = 0;
  bool mlp_ready = false;
  bool feat_corpus_ready = false;       // wnorm / dstat / dcdf for the current corpus
  DevBuf<float> wnorm, dstat, dcdf;     // corpus features
  DevBuf<float> qstat, qc, qn;          // query features of the current batch
  DevBuf<float> pred;                   // predictions
  double last_lamda = 0.0;              // lamda of the last run (its mrow)
  DevBuf<float> duals;
  bool duals_valid = false;
  DevBuf<uint32_t> fallbacks;
  DevBuf<float> warm_vals;              // caller-supplied starting log v
  bool warm_given = false;
};

namespace {

int check_ctx(wmd_context* ctx) {
  if (!ctx) {
    set_err("null context");
    return WMD_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  return WMD_OK;
}

// Fills SolveArgs fields common to every variant.
wmd::SolveArgs base_args(wmd_context* c, double lamda, const wmd_options& o) {
  wmd::SolveArgs a{};
  std::memset(&a, 0, sizeof(a));  // padding too: cached graphs compare arguments bytewise
  a.K = c->K.p;
  a.mrow = c->mrow.p;
  a.k_off = c->k_off.p;
  a.k_stride = c->k_stride.p;
  a.qmeta = c->qmeta.p;
  a.q_ptr = c->q_ptr.p;
  a.q_r = c->q_r.p;
  a.doc_ptr = c->doc_ptr.p;
  a.doc_t = c->doc_t.p;
  a.doc_t1 = c->doc_t1.p;
  a.doc_val = c->doc_val.p;
  a.nnz = c->nnz;
  a.N = (int)c->N;
  a.Vt = (int)c->Vt;
  a.max_iter = o.max_iter;
  a.tol = (float)o.tol;
  a.omega = (float)o.omega;
  a.accel = o.accel;
  a.inv_lamda = (float)(1.0 / lamda);
  a.lamda = (float)lamda;
  a.wmd = c->wmd.p;
  a.iters = c->iters.p;
  a.resid = c->resid.p;
  a.lds_q = round_up4(std::max(c->qmax, 1));
  a.lds_n = round_up4(std::max(c->nmax, 1));
  return a;
}

// The resident solver addresses a query's kernel block with 32-bit byte
// offsets (buffer loads) and forms row offsets with 24-bit multiplies: every
// block must stay below 2 GiB and the corpus below 2^24 distinct words.
bool resident_blocks_fit(const wmd_context* c) {
  if (c->Vt >= (int64_t)1 << 24) return false;
  for (int Q = 0; Q < c->nq; ++Q) {
    const int64_t ks = round_up4(c->h_q_ptr[Q + 1] - c->h_q_ptr[Q]);
    if (c->Vt * ks * 4 >= ((int64_t)1 << 31)) return false;
  }
  return true;
}

// Whether shape s has a kernel for the domain.
bool shape_in_domain(int s, bool log_domain) {
  return log_domain ? wmd::kShapes[s].log != 0 : wmd::kShapes[s].lin != 0;
}

// Smallest-capacity tile shape of the domain that holds q query words and
// documents of nmax words.
int pick_shape(int q, int nmax, bool log_domain) {
  for (int s = 0; s < wmd::kNumShapes; ++s)
    if (wmd::kShapes[s].cap_q() >= q && wmd::kShapes[s].cap_n() >= nmax &&
        shape_in_domain(s, log_domain))
      return s;
  return -1;
}

bool shape_fits(int s, int q, bool log_domain) {
  return s >= 0 && s < wmd::kNumShapes && wmd::kShapes[s].cap_q() >= q &&
         shape_in_domain(s, log_domain);
}

// Tile shape for a query of q words against documents of class cls: the
// forced shape, else the tuned table's entry, else the smallest that fits.
int shape_for(const wmd_context* c, int cls, int q, bool log_domain, int forced) {
  if (forced >= 0) return forced;
  const std::vector<int32_t>& t = c->tune[log_domain ? 1 : 0];
  if (!t.empty()) {
    const int s = t[(size_t)cls * kTuneStride + q];
    if (shape_fits(s, q, log_domain)) return s;
  }
  return pick_shape(q, 0, log_domain);
}

// Until a run has measured them, every pair is weighted with this many
// iterations.
constexpr double kPriorIters = 20.0;
// Largest share of a plan's estimated work up to which consecutive classes
// merge into one launch.
constexpr double kMergeShare = 0.004;

void add_launch(std::vector<ResidentLaunch>* plan, std::vector<int32_t>* flat,
                const std::vector<int32_t>& queries, int s, int cls, int64_t d0, int64_t d1) {
  if (queries.empty() || d1 == d0) return;
  ResidentLaunch l{s, cls, flat->size(), (int64_t)queries.size(), d0, d1, 0.0};
  l.work = (double)l.cnt * (double)(d1 - d0) * kPriorIters * iter_cost(s, cls);
  flat->insert(flat->end(), queries.begin(), queries.end());
  plan->push_back(l);
}

// Per-iteration estimate of a long-document pair (q query words, n document
// words) in the units of iter_cost (wave instructions): two passes over the
// q x n kernel entries by 64-lane waves.
double long_iter_cost(int q, int n) { return 2.0 * (double)q * (double)n / 64.0; }

// Every (document class, tile shape) group of the current queries; the
// bucket array receives the query ids of each group (query-major pairs, so
// concurrently running waves share one query's kernel rows in cache).
int plan_resident(wmd_context* c, int forced_shape, bool log_domain,
                  std::vector<ResidentLaunch>* plan) {
  std::vector<int32_t> flat;
  plan->clear();
  // groups[cls][s]: the queries that run shape s on document class cls.
  std::vector<std::vector<std::vector<int32_t>>> groups(
      wmd::kNumDocClasses, std::vector<std::vector<int32_t>>(wmd::kNumShapes));
  for (int cls = 0; cls < wmd::kNumDocClasses; ++cls) {
    if (c->h_class_off[cls + 1] == c->h_class_off[cls]) continue;
    for (int Q = 0; Q < c->nq; ++Q) {
      const int q = c->h_q_ptr[Q + 1] - c->h_q_ptr[Q];
      groups[cls][shape_for(c, cls, q, log_domain, forced_shape)].push_back(Q);
    }
  }
  // Consecutive classes whose tiles have the same row count NE and whose
  // groups hold the same queries form one launch (their documents are
  // contiguous in doc_perm) while the merged launch stays small
  // (kMergeShare of the estimated work): a small launch's time is its
  // slowest pair's latency, so fewer of them shorten the lanes' tails, while
  // large launches stay apart for the lane balance.  The launch takes the
  // last class's NE, which is the same.
  auto group_work = [&](int cls, int s) {
    return (double)groups[cls][s].size() *
           (double)(c->h_class_off[cls + 1] - c->h_class_off[cls]) * kPriorIters *
           iter_cost(s, cls);
  };
  double total_est = 0.0;
  for (int cls = 0; cls < wmd::kNumDocClasses; ++cls)
    for (int s = 0; s < wmd::kNumShapes; ++s) total_est += group_work(cls, s);
  for (int s = 0; s < wmd::kNumShapes; ++s) {
    for (int cls = 0; cls < wmd::kNumDocClasses;) {
      int last = cls;
      if (!groups[cls][s].empty()) {
        double w = group_work(cls, s);
        while (last + 1 < wmd::kNumDocClasses && tile_rows(s, last + 1) == tile_rows(s, cls) &&
               groups[last + 1][s] == groups[cls][s] &&
               w + group_work(last + 1, s) <= kMergeShare * total_est) {
          ++last;
          w += group_work(last, s);
        }
      }
      add_launch(plan, &flat, groups[cls][s], s, last, c->h_class_off[cls],
                 c->h_class_off[last + 1]);
      cls = last + 1;
    }
  }
  if (c->n_long > 0) {
    // Every query against the documents beyond the tiles, in one launch.
    ResidentLaunch l{kLongLaunch, -1, flat.size(), (int64_t)c->nq, 0, c->n_long, 0.0};
    for (int Q = 0; Q < c->nq; ++Q) {
      flat.push_back(Q);
      const int q = c->h_q_ptr[Q + 1] - c->h_q_ptr[Q];
      for (int32_t j : c->h_long)
        l.work += 4.0 * kPriorIters * long_iter_cost(q, c->h_doc_ptr[j + 1] - c->h_doc_ptr[j]);
    }
    plan->push_back(l);
  }
  HIP_TRY(c->plan_bucket.upload(flat.data(), flat.size(), c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->h_plan_flat = std::move(flat);
  c->plan_measured = false;
  return WMD_OK;
}

// Re-weights the cached plan with the iterations its pairs took in the run
// that just finished (it: nq x N counts) and bumps the plan id, so the next
// run (and its captured graph) balances the lanes on measured work.  Done
// once per plan; the launch order does not change any result.
void measure_plan(wmd_context* c, const std::vector<int32_t>& it) {
  if (c->plan_measured || c->plan.empty() || c->h_doc_perm.empty()) return;
  for (ResidentLaunch& l : c->plan) {
    if (l.shape == kLongLaunch) {
      double w = 0.0;
      for (int64_t k = 0; k < l.cnt; ++k) {
        const int Q = c->h_plan_flat[l.off + k];
        const int q = c->h_q_ptr[Q + 1] - c->h_q_ptr[Q];
        for (int64_t d = l.d0; d < l.d1; ++d) {
          const int32_t j = c->h_long[d];
          w += (double)it[(size_t)Q * c->N + j] *
               long_iter_cost(q, c->h_doc_ptr[j + 1] - c->h_doc_ptr[j]);
        }
      }
      l.work = w;
      continue;
    }
    double iters = 0.0;
    for (int64_t k = 0; k < l.cnt; ++k) {
      const int32_t* row = it.data() + (size_t)c->h_plan_flat[l.off + k] * c->N;
      for (int64_t d = l.d0; d < l.d1; ++d) iters += row[c->h_doc_perm[d]];
    }
    l.work = iters * iter_cost(l.shape, l.cls);
  }
  c->plan_measured = true;
  ++c->plan_id;
}

// Launch l with the query ids at bucket + l.off.
hipError_t launch_one(wmd_context* c, const wmd::SolveArgs& base, const ResidentLaunch& l,
                      const int32_t* bucket, bool log_domain, hipStream_t s) {
  wmd::SolveArgs a = base;
  a.bucket_q = bucket + l.off;
  a.n_list = (int)(l.d1 - l.d0);
  a.n_pairs = l.cnt * (l.d1 - l.d0);
  if (l.shape == kLongLaunch) {
    a.doc_list = c->long_docs.p + l.d0;
    return wmd::launch_long(a, log_domain, s);
  }
  a.doc_list = c->doc_perm.p + l.d0;
  return wmd::launch_resident(l.shape, l.cls, log_domain, a, s);
}

// Launches a plan on the four lanes: largest first, each to the lane with
// the least queued work (longest-processing-time list scheduling), so a
// launch's tail of slow pairs overlaps other launches.  The side streams
// fork from the context stream before and join it after.
int launch_plan(wmd_context* c, const wmd::SolveArgs& base,
                const std::vector<ResidentLaunch>& plan, bool log_domain, int* launches) {
  // The long-document launch (few blocks, each iterating in the kernel)
  // goes first so its blocks overlap the whole stage; the rest largest first.
  std::vector<size_t> order(plan.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
    const bool lx = plan[x].shape == kLongLaunch, ly = plan[y].shape == kLongLaunch;
    if (lx != ly) return lx;
    return plan[x].work > plan[y].work;
  });
  HIP_TRY(hipEventRecord(c->fork, c->stream));
  for (int k = 0; k < kSideStreams; ++k) HIP_TRY(hipStreamWaitEvent(c->side[k], c->fork, 0));
  double load[kLanes] = {};
  for (size_t i : order) {
    const int k = (int)(std::min_element(load, load + kLanes) - load);
    load[k] += plan[i].work;
    HIP_TRY(launch_one(c, base, plan[i], c->plan_bucket.p, log_domain,
                       k == 0 ? c->stream : c->side[k - 1]));
    ++*launches;
  }
  for (int k = 0; k < kSideStreams; ++k) {
    HIP_TRY(hipEventRecord(c->join[k], c->side[k]));
    HIP_TRY(hipStreamWaitEvent(c->stream, c->join[k], 0));
  }
  return WMD_OK;
}

void drop_solve_graph(wmd_context* c) {
  if (c->solve_exec) (void)hipGraphExecDestroy(c->solve_exec);
  c->solve_exec = nullptr;
}

// The resident solver stage.  The plan (and its bucket upload) is rebuilt
// only when queries, corpus, tuning or the shape/domain choice changed.  With
// use_graph the stage's launches, fork and join included, are captured once
// into a HIP graph and replayed while the plan and the kernel arguments stay
// the same (the arguments are compared bytewise; base_args zero-fills them).
int run_resident(wmd_context* c, const wmd::SolveArgs& base, int forced_shape, bool log_domain,
                 bool use_graph, int* launches) {
  if (c->plan_built != c->plan_epoch || c->plan_forced != forced_shape ||
      c->plan_log != log_domain) {
    int rc = plan_resident(c, forced_shape, log_domain, &c->plan);
    if (rc != WMD_OK) return rc;
    c->plan_built = c->plan_epoch;
    c->plan_forced = forced_shape;
    c->plan_log = log_domain;
    ++c->plan_id;
  }
  if (!use_graph) return launch_plan(c, base, c->plan, log_domain, launches);
  if (!c->solve_exec || c->solve_plan_id != c->plan_id ||
      std::memcmp(&c->solve_args, &base, sizeof(base)) != 0) {
    drop_solve_graph(c);
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int captured = 0;
    const int rc = launch_plan(c, base, c->plan, log_domain, &captured);
    hipGraph_t graph = nullptr;
    const hipError_t end = hipStreamEndCapture(c->stream, &graph);
    if (rc != WMD_OK || end != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      if (rc == WMD_OK) set_err("solver-stage capture failed: %s", hipGetErrorString(end));
      return rc != WMD_OK ? rc : WMD_ERR_HIP;
    }
    const hipError_t ie = hipGraphInstantiate(&c->solve_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      c->solve_exec = nullptr;
      set_err("solver-stage graph instantiation failed: %s", hipGetErrorString(ie));
      return WMD_ERR_HIP;
    }
    c->solve_args = base;
    c->solve_plan_id = c->plan_id;
  }
  HIP_TRY(hipGraphLaunch(c->solve_exec, c->stream));
  *launches += (int)c->plan.size();
  return WMD_OK;
}

// Streaming variants: iterations launched from a captured HIP graph.
int run_streaming(wmd_context* c, wmd::SolveArgs a, bool fused, const wmd_options& o,
                  int* launches) {
  std::vector<int32_t> ident(c->nq);
  for (int Q = 0; Q < c->nq; ++Q) ident[Q] = Q;
  HIP_TRY(c->bucket.upload(ident.data(), ident.size(), c->stream));
  HIP_TRY(c->u.ensure((size_t)c->u_total));
  HIP_TRY(c->done.ensure((size_t)c->nq * c->N));
  HIP_TRY(c->status.ensure(2));
  if (!fused) HIP_TRY(c->ow.ensure((size_t)c->nq * (size_t)c->nnz));
  a.bucket_q = c->bucket.p;
  a.n_pairs = (int64_t)c->nq * c->N;
  if (a.n_pairs > INT32_MAX) {  // pair indices (and the active list) are 32-bit
    set_err("%lld (query, document) pairs exceed the streaming launch's 2^31 - 1 pairs",
            (long long)a.n_pairs);
    return WMD_ERR_UNSUPPORTED;
  }
  a.u = c->u.p;
  a.u_off = c->u_off.p;
  a.done = c->done.p;
  a.ow = c->ow.p;
  auto step = [&](void) -> hipError_t {
    if (fused) return wmd::launch_fused_step(a, false, c->stream);
    hipError_t e = wmd::launch_sddmm(a, c->stream);
    if (e != hipSuccess) return e;
    return wmd::launch_spmm(a, false, c->stream);
  };
  const int per_step = fused ? 1 : 2;
  int done_iters = 0;
  if (o.max_iter > 0 && !a.log_domain) {
    // The first iteration starts from the uniform u itself and initializes
    // the per-pair state (no separate initialization launch).
    a.first = 1;
    HIP_TRY(step());
    a.first = 0;
    *launches += per_step;
    done_iters = 1;
  } else {
    HIP_TRY(wmd::launch_stream_init(a, c->stream));
    ++*launches;
  }
  // With a tolerance, every check_every iterations a status scan lists the
  // pairs still iterating (the iterations walk that list); the host reads the
  // list's length after every kScansPerCheck scans.  Lists and counts are
  // double-buffered: the iterations of chunk k read list/count k % 2 and zero
  // count (k + 1) % 2, which the chunk's scan then fills.
  constexpr int kScansPerCheck = 8;
  const bool use_tol = o.tol > 0.0;
  const int chunk = use_tol ? std::max(1, o.check_every > 0 ? o.check_every : 8) : o.max_iter;
  const wmd::SolveArgs all_pairs = a;  // the final pass covers every pair
  int32_t* list[2] = {nullptr, nullptr};
  uint32_t* count[2] = {c->status.p, c->status.p + 1};
  if (use_tol && done_iters < o.max_iter) {
    HIP_TRY(c->act.ensure(2 * (size_t)a.n_pairs));
    list[0] = c->act.p;
    list[1] = c->act.p + a.n_pairs;
    HIP_TRY(wmd::launch_zero_u32(count[0], 1, c->stream));
    HIP_TRY(wmd::launch_scan(a, nullptr, nullptr, list[0], count[0], c->stream));
    *launches += 2;
  }
  const int span = use_tol ? chunk * kScansPerCheck : chunk;  // iterations per graph
  int rc = WMD_OK;
  while (done_iters < o.max_iter) {
    const int len = std::min(span, o.max_iter - done_iters);
    const int n_chunks = use_tol ? (len + chunk - 1) / chunk : 1;
    if (!c->stream_exec || c->stream_len != len || c->stream_fused != fused ||
        c->stream_tol != use_tol || c->stream_chunk != chunk ||
        std::memcmp(&c->stream_args, &a, sizeof(a)) != 0) {
      if (c->stream_exec) (void)hipGraphExecDestroy(c->stream_exec);
      c->stream_exec = nullptr;
      hipGraph_t graph = nullptr;
      HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      hipError_t cap = hipSuccess;
      const wmd::SolveArgs a0 = a;
      for (int k = 0, it = 0; k < n_chunks && cap == hipSuccess; ++k) {
        const int steps = use_tol ? std::min(chunk, len - it) : len;
        if (use_tol) {
          a.act = list[k & 1];
          a.n_act = count[k & 1];
          a.n_clear = count[(k + 1) & 1];
        }
        for (int t = 0; t < steps && cap == hipSuccess; ++t) cap = step();
        if (cap == hipSuccess && use_tol)
          cap = wmd::launch_scan(a, list[k & 1], count[k & 1], list[(k + 1) & 1],
                                 count[(k + 1) & 1], c->stream);
        it += steps;
      }
      a = a0;
      hipError_t end = hipStreamEndCapture(c->stream, &graph);
      if (cap != hipSuccess || end != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        set_err("graph capture failed: %s", hipGetErrorString(cap != hipSuccess ? cap : end));
        return WMD_ERR_HIP;
      }
      hipError_t ie = hipGraphInstantiate(&c->stream_exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        c->stream_exec = nullptr;
        set_err("hipGraphInstantiate failed: %s", hipGetErrorString(ie));
        return WMD_ERR_HIP;
      }
      c->stream_args = a;
      c->stream_len = len;
      c->stream_fused = fused;
      c->stream_tol = use_tol;
      c->stream_chunk = chunk;
    }
    hipGraphExec_t exec = c->stream_exec;
    hipError_t le = hipGraphLaunch(exec, c->stream);
    if (le != hipSuccess) {
      set_err("hipGraphLaunch failed: %s", hipGetErrorString(le));
      rc = WMD_ERR_HIP;
      break;
    }
    *launches += len * per_step + (use_tol ? n_chunks : 0);
    done_iters += len;
    if (use_tol) {
      // A graph of an even number of chunks leaves the list in slot 0.
      uint32_t* last = count[n_chunks & 1];
      hipError_t ce = hipMemcpyAsync(c->h_status, last, sizeof(uint32_t),
                                     hipMemcpyDeviceToHost, c->stream);
      if (ce == hipSuccess) ce = hipStreamSynchronize(c->stream);
      if (ce != hipSuccess) {
        set_err("status readback failed: %s", hipGetErrorString(ce));
        rc = WMD_ERR_HIP;
        break;
      }
      if (c->h_status[0] == 0) break;  // every pair has converged
      if (n_chunks & 1) break;  // an odd (last, partial) graph: max_iter reached
    }
  }
  if (rc != WMD_OK) return rc;
  a = all_pairs;
  if (fused) {
    HIP_TRY(wmd::launch_fused_step(a, true, c->stream));
    ++*launches;
  } else {
    // Final SDDMM ignores the stop flags: clear them first (the converged
    // pairs' w must be recomputed from their final u).
    HIP_TRY(hipMemsetAsync(c->done.p, 0, (size_t)c->nq * c->N, c->stream));
    HIP_TRY(wmd::launch_sddmm(a, c->stream));
    HIP_TRY(wmd::launch_spmm(a, true, c->stream));
    *launches += 2;
  }
  return WMD_OK;
}

int validate_csr(uint64_t V, uint64_t N, const uint64_t* row_ptr, const uint64_t* col_inds,
                 const double* vals) {
  if (!row_ptr || V == 0 || N == 0) {
    set_err("empty corpus or null row_ptr");
    return WMD_ERR_ARG;
  }
  if (V >= (1ull << 31) || N >= (1ull << 31)) {
    set_err("V and N must be < 2^31");
    return WMD_ERR_UNSUPPORTED;
  }
  if (row_ptr[0] != 0) {
    set_err("row_ptr[0] != 0");
    return WMD_ERR_ARG;
  }
  for (uint64_t w = 0; w < V; ++w)
    if (row_ptr[w + 1] < row_ptr[w]) {
      set_err("row_ptr decreases at row %llu", (unsigned long long)w);
      return WMD_ERR_ARG;
    }
  const uint64_t nnz = row_ptr[V];
  if (nnz >= (1ull << 31)) {
    set_err("nnz must be < 2^31");
    return WMD_ERR_UNSUPPORTED;
  }
  if (nnz > 0 && (!col_inds || !vals)) {
    set_err("null col_inds/vals");
    return WMD_ERR_ARG;
  }
  for (uint64_t e = 0; e < nnz; ++e) {
    if (col_inds[e] >= N) {
      set_err("col_inds[%llu] = %llu out of range", (unsigned long long)e,
              (unsigned long long)col_inds[e]);
      return WMD_ERR_ARG;
    }
    if (!(vals[e] > 0.0) || !std::isfinite(vals[e])) {
      set_err("vals[%llu] must be finite and > 0", (unsigned long long)e);
      return WMD_ERR_ARG;
    }
  }
  return WMD_OK;
}

int one_shot(const double* r, const double* vecs, double* WMD, double lamda, uint64_t V,
             uint64_t N, uint64_t D, const uint64_t* row_ptr, const uint64_t* col_inds,
             const double* vals, const wmd_options* opt, wmd_stats* stats) {
  if (!r || !vecs || !WMD) {
    set_err("null r/vecs/WMD");
    return WMD_ERR_ARG;
  }
  // select_non_empty_entries (sinkhorn_wmd_fused.cpp:45-63), ascending word order.
  std::vector<int64_t> qp{0}, qw;
  std::vector<double> qv;
  for (uint64_t w = 0; w < V; ++w) {
    if (!(r[w] >= 0.0)) {
      set_err("r[%llu] is negative or NaN", (unsigned long long)w);
      return WMD_ERR_ARG;
    }
    if (r[w] > 0.0) {
      qw.push_back((int64_t)w);
      qv.push_back(r[w]);
    }
  }
  if (qw.empty()) {
    set_err("r has no positive entry");
    return WMD_ERR_ARG;
  }
  qp.push_back((int64_t)qw.size());
  wmd_context* ctx = nullptr;
  int rc = wmd_context_create(0, &ctx);
  if (rc == WMD_OK)
    rc = wmd_context_set_corpus(ctx, nullptr, vecs, V, D, row_ptr, col_inds, vals, N);
  if (rc == WMD_OK) rc = wmd_context_set_queries(ctx, 1, qp.data(), qw.data(), qv.data());
  if (rc == WMD_OK) rc = wmd_context_run(ctx, lamda, opt, stats);
  if (rc == WMD_OK) rc = wmd_context_result(ctx, WMD, nullptr, nullptr);
  std::string keep = g_err;
  wmd_context_destroy(ctx);
  g_err = keep;
  return rc;
}

}  // namespace

// ---------------------------------------------------------------------------
// Learned warm start: features, inference stage.
namespace {

// Word norms, document statistics and mid-ranks: once per corpus, outside the
// timed stages (like the matrix-core planes).
int warm_corpus_features(wmd_context* c) {
  if (c->feat_corpus_ready) return WMD_OK;
  HIP_TRY(c->wnorm.ensure((size_t)c->Vt));
  HIP_TRY(c->dstat.ensure((size_t)c->N * 5));
  HIP_TRY(c->dcdf.ensure((size_t)c->nnz));
  HIP_TRY(wmd::launch_word_norms(c->E.p, c->touched.p, (int)c->Vt, (int)c->D, c->wnorm.p,
                                 c->stream));
  HIP_TRY(wmd::launch_doc_feats(c->doc_ptr.p, c->doc_t.p, c->doc_val.p, c->wnorm.p, (int)c->N,
                                c->dstat.p, c->dcdf.p, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->feat_corpus_ready = true;
  return WMD_OK;
}

// Query statistics and sorted norm tables of the current batch (every run:
// they belong to the queries, so a serving loop pays them per batch).
int warm_query_features(wmd_context* c, int* launches) {
  if (c->qmax > wmd::kWarmQueryMax) {
    set_err("the learned warm start takes queries of at most %d words (longest %d)",
            wmd::kWarmQueryMax, c->qmax);
    return WMD_ERR_UNSUPPORTED;
  }
  const size_t sq = (size_t)c->h_q_ptr[c->nq];
  HIP_TRY(c->qstat.ensure((size_t)c->nq * 5));
  HIP_TRY(c->qc.ensure(sq));
  HIP_TRY(c->qn.ensure(sq));
  HIP_TRY(wmd::launch_query_feats(c->q_ptr.p, c->q_words.p, c->q_r.p, c->E.p, (int)c->D, c->nq,
                                  c->qstat.p, c->qc.p, c->qn.p, c->stream));
  ++*launches;
  return WMD_OK;
}

wmd::WarmInferArgs warm_args(wmd_context* c, double lamda) {
  wmd::WarmInferArgs a{};
  a.c.wnorm = c->wnorm.p;
  a.c.dstat = c->dstat.p;
  a.c.dcdf = c->dcdf.p;
  a.c.doc_t = c->doc_t.p;
  a.c.doc_of = c->doc_of.p;
  a.c.nnz = c->nnz;
  a.q.qstat = c->qstat.p;
  a.q.qc = c->qc.p;
  a.q.qn = c->qn.p;
  a.q.q_ptr = c->q_ptr.p;
  a.q.nq = c->nq;
  a.q.qmax = c->qmax;
  a.mrow = c->mrow.p;
  a.Vt = (int)c->Vt;
  a.lam_abs = (float)(-lamda);
  a.params = c->mlp_p.p;
  return a;
}

// Inference stage (after the distance stage, which leaves mrow): query
// features, then every (query, nonzero) sample into c->pred.
int mlp_predict(wmd_context* c, double lamda, int* launches) {
  int rc = warm_query_features(c, launches);
  if (rc != WMD_OK) return rc;
  HIP_TRY(c->pred.ensure((size_t)c->nq * (size_t)c->nnz));
  wmd::WarmInferArgs a = warm_args(c, lamda);
  a.pred = c->pred.p;
  HIP_TRY(wmd::launch_warm_infer(a, c->stream));
  *launches += 1;
  return WMD_OK;
}

// Host-side checks of a training batch before its launch: every sample
// names a query and a nonzero of the context, targets index the dual buffer,
// and the group offsets run from 0 to the batch size without decreasing.
int check_train_batch(const wmd_context* c, const int32_t* sq, const int32_t* spos,
                      const int64_t* yidx, int64_t nb, const int32_t* rel, int groups) {
  if (rel[0] != 0 || rel[groups] != nb) {
    set_err("training batch offsets do not cover the batch");
    return WMD_ERR_STATE;
  }
  for (int g = 0; g < groups; ++g)
    if (rel[g + 1] < rel[g]) {
      set_err("training batch offsets decrease");
      return WMD_ERR_STATE;
    }
  const int64_t ndual = (int64_t)c->nq * c->nnz;
  for (int64_t s = 0; s < nb; ++s)
    if (sq[s] < 0 || sq[s] >= c->nq || spos[s] < 0 || spos[s] >= c->nnz || yidx[s] < 0 ||
        yidx[s] >= ndual) {
      set_err("training sample %lld out of range", (long long)s);
      return WMD_ERR_STATE;
    }
  return WMD_OK;
}

}  // namespace


extern "C" {

void wmd_options_default(wmd_options* opt) {
  if (!opt) return;
  std::memset(opt, 0, sizeof(*opt));
  opt->max_iter = 15;
  opt->variant = WMD_VARIANT_RESIDENT;
  opt->tol = 0.0;
  opt->check_every = 8;
  opt->shape = -1;
  opt->summarize = 1;
  opt->domain = WMD_DOMAIN_LINEAR;
  opt->cdist = WMD_CDIST_MFMA;
  opt->warm_start = WMD_WARM_UNIFORM;
  opt->export_duals = 0;
  opt->accel = WMD_ACCEL_NONE;
  opt->omega = 1.0;
  opt->graph = 1;
}

const char* wmd_last_error(void) { return g_err.c_str(); }

int wmd_num_shapes(void) { return wmd::kNumShapes; }

int wmd_shape_info(int shape, int32_t* waves, int32_t* cap_q, int32_t* cap_n) {
  if (shape < 0 || shape >= wmd::kNumShapes) {
    set_err("shape %d out of range", shape);
    return WMD_ERR_ARG;
  }
  if (waves) *waves = wmd::kShapes[shape].W;
  if (cap_q) *cap_q = wmd::kShapes[shape].cap_q();
  if (cap_n) *cap_n = wmd::kShapes[shape].cap_n();
  return WMD_OK;
}

int wmd_shape_domains(int shape) {
  if (shape < 0 || shape >= wmd::kNumShapes) {
    set_err("shape %d out of range", shape);
    return WMD_ERR_ARG;
  }
  return (wmd::kShapes[shape].lin ? 1 : 0) | (wmd::kShapes[shape].log ? 2 : 0);
}

int wmd_context_create(int device, wmd_context** out) {
  if (!out) {
    set_err("null output pointer");
    return WMD_ERR_ARG;
  }
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) {
    set_err("device %d not present (%d devices)", device, ndev);
    return WMD_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(device));
  unsigned char* scr = nullptr;
  int rc = scratch_address(device, &scr);
  if (rc != WMD_OK) return rc;
  wmd_context* c = new wmd_context();
  c->device = device;
  c->scratch = scr;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipEventCreate(&c->ev[i]);
  for (int i = 0; i < kSideStreams && e == hipSuccess; ++i) {
    e = hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join[i], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&c->h_status), 4 * sizeof(uint32_t), hipHostMallocDefault);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) {
    set_err("context setup failed: %s", hipGetErrorString(e));
    wmd_context_destroy(c);
    return WMD_ERR_HIP;
  }
  *out = c;
  return WMD_OK;
}

void wmd_context_destroy(wmd_context* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->E.release(); c->touched.release(); c->doc_ptr.release(); c->doc_t.release();
  c->doc_t1.release(); c->doc_val.release(); c->q_ptr.release(); c->q_words.release(); c->k_stride.release();
  c->qmeta.release(); c->act.release();
  c->bucket.release(); c->q_r.release(); c->k_off.release(); c->u_off.release();
  c->K.release(); c->mrow.release(); c->wmd.release(); c->iters.release();
  c->resid.release(); c->u.release(); c->ow.release(); c->done.release();
  c->doc_perm.release();
  c->long_docs.release();
  c->Es.release(); c->nt.release(); c->Qs.release(); c->nqw.release();
  c->doc_of.release(); c->mlp_p.release(); c->mlp_m.release(); c->mlp_v.release();
  c->wnorm.release(); c->dstat.release(); c->dcdf.release(); c->qstat.release();
  c->qc.release(); c->qn.release(); c->pred.release();
  c->duals.release();
  c->fallbacks.release(); c->warm_vals.release();
  c->status.release();
  c->plan_bucket.release();
  c->d_words.release(); c->slot.release(); c->dq_ptr.release(); c->dk_stride.release();
  c->dk_off.release(); c->Dd.release();
  drop_solve_graph(c);
  if (c->stream_exec) (void)hipGraphExecDestroy(c->stream_exec);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (int i = 0; i < kSideStreams; ++i) {
    if (c->side[i]) (void)hipStreamSynchronize(c->side[i]);
    if (c->side[i]) (void)hipStreamDestroy(c->side[i]);
    if (c->join[i]) (void)hipEventDestroy(c->join[i]);
  }
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->h_status) (void)hipHostFree(c->h_status);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int wmd_context_set_corpus(wmd_context* c, const float* vecs_f32, const double* vecs_f64,
                           uint64_t V, uint64_t D, const uint64_t* row_ptr,
                           const uint64_t* col_inds, const double* vals, uint64_t N) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if ((vecs_f32 == nullptr) == (vecs_f64 == nullptr)) {
    set_err("exactly one of vecs_f32 / vecs_f64 must be given");
    return WMD_ERR_ARG;
  }
  if (D == 0 || D > 65536) {
    set_err("embedding size D=%llu unsupported", (unsigned long long)D);
    return WMD_ERR_ARG;
  }
  rc = validate_csr(V, N, row_ptr, col_inds, vals);
  if (rc != WMD_OK) return rc;
  const uint64_t nnz = row_ptr[V];
  // Compact word ids of words that occur, and the doc-major CSR over them.
  std::vector<int32_t> compact(V, -1), touched;
  for (uint64_t w = 0; w < V; ++w)
    if (row_ptr[w + 1] > row_ptr[w]) {
      compact[w] = (int32_t)touched.size();
      touched.push_back((int32_t)w);
    }
  std::vector<int32_t> dptr(N + 1, 0), fill(N, 0), dt(nnz);
  std::vector<float> dv(nnz);
  for (uint64_t e = 0; e < nnz; ++e) dptr[col_inds[e] + 1]++;
  int nmax = 0;
  for (uint64_t j = 0; j < N; ++j) {
    nmax = std::max(nmax, dptr[j + 1]);
    dptr[j + 1] += dptr[j];
  }
  for (uint64_t w = 0; w < V; ++w)
    for (uint64_t e = row_ptr[w]; e < row_ptr[w + 1]; ++e) {
      const uint64_t j = col_inds[e];
      const int32_t pos = dptr[j] + fill[j]++;
      dt[pos] = compact[w];
      dv[pos] = (float)vals[e];
    }
  // The embeddings go to the device as given (FP64 input into a temporary);
  // the FP32 conversion and the finiteness check run there, so the host does
  // no pass over V x D.  From here on a failure leaves the context without a
  // corpus.
  c->has_corpus = false;
  c->has_queries = false;
  ScopedDev e64, flag;
  HIP_TRY(c->E.ensure(V * D));
  HIP_TRY(hipMalloc(&flag.p, sizeof(unsigned long long)));
  if (vecs_f64) {
    HIP_TRY(hipMalloc(&e64.p, V * D * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(e64.p, vecs_f64, V * D * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
  } else {
    HIP_TRY(hipMemcpyAsync(c->E.p, vecs_f32, V * D * sizeof(float), hipMemcpyHostToDevice,
                           c->stream));
  }
  HIP_TRY(wmd::launch_embed(static_cast<const double*>(e64.p), c->E.p, V * D,
                            static_cast<unsigned long long*>(flag.p), c->stream));
  unsigned long long first_bad = 0;
  HIP_TRY(hipMemcpyAsync(&first_bad, flag.p, sizeof(first_bad), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (first_bad != ~0ull) {
    set_err("embedding element %llu is not finite", first_bad);
    return WMD_ERR_ARG;
  }
  HIP_TRY(c->touched.upload(touched.data(), touched.size(), c->stream));
  HIP_TRY(c->doc_ptr.upload(dptr.data(), dptr.size(), c->stream));
  HIP_TRY(c->doc_t.upload(dt.data(), dt.size(), c->stream));
  {
    std::vector<int32_t> dt1(dt.size());
    for (size_t i = 0; i < dt.size(); ++i) dt1[i] = dt[i] + 1;
    HIP_TRY(c->doc_t1.upload(dt1.data(), dt1.size(), c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  HIP_TRY(c->doc_val.upload(dv.data(), dv.size(), c->stream));
  // Documents grouped by length class (stable, so ids ascend inside a
  // class); documents beyond the last class go to the streaming kernel in
  // resident runs (long_docs, ascending).
  std::vector<int64_t> class_off(wmd::kNumDocClasses + 1, 0);
  std::vector<int32_t> perm;
  perm.reserve(N);
  for (int cls = 0; cls < wmd::kNumDocClasses; ++cls) {
    class_off[cls] = (int64_t)perm.size();
    for (uint64_t j = 0; j < N; ++j) {
      const int len = dptr[j + 1] - dptr[j];
      const int cj = len == 0 ? 0 : (len - 1) / wmd::kDocClassWidth;
      if (cj == cls) perm.push_back((int32_t)j);
    }
  }
  class_off[wmd::kNumDocClasses] = (int64_t)perm.size();
  HIP_TRY(c->doc_perm.upload(perm.data(), perm.size(), c->stream));
  std::vector<int32_t> long_docs;
  for (uint64_t j = 0; j < N; ++j)
    if (dptr[j + 1] - dptr[j] > wmd::kShapes[0].cap_n()) long_docs.push_back((int32_t)j);
  HIP_TRY(c->long_docs.upload(long_docs.data(), long_docs.size(), c->stream));
  std::vector<int32_t> doc_of(nnz);
  for (uint64_t j = 0; j < N; ++j)
    for (int32_t pos = dptr[j]; pos < dptr[j + 1]; ++pos) doc_of[pos] = (int32_t)j;
  HIP_TRY(c->doc_of.upload(doc_of.data(), doc_of.size(), c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->h_class_off = class_off;
  c->h_doc_perm = perm;
  c->h_doc_ptr = dptr;
  c->h_doc_t = dt;
  c->h_touched = touched;
  c->feat_corpus_ready = false;
  c->duals_valid = false;
  c->V = (int64_t)V;
  c->D = (int64_t)D;
  c->N = (int64_t)N;
  c->nnz = (int64_t)nnz;
  c->Vt = (int64_t)touched.size();
  c->nmax = nmax;
  c->n_long = (int)long_docs.size();
  c->h_long = long_docs;
  c->split_ready = false;
  c->has_corpus = true;
  ++c->plan_epoch;
  c->has_queries = false;  // query blocks depend on Vt and N
  c->ran = false;
  return WMD_OK;
}

int wmd_context_set_queries(wmd_context* c, uint64_t n, const int64_t* qp, const int64_t* qw,
                            const double* qv) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if (!c->has_corpus) {
    set_err("set_corpus must be called before set_queries");
    return WMD_ERR_STATE;
  }
  if (n == 0 || n >= (1u << 20) || !qp || !qw || !qv) {
    set_err("bad query batch (n=%llu)", (unsigned long long)n);
    return WMD_ERR_ARG;
  }
  if (qp[0] != 0) {
    set_err("q_ptr[0] != 0");
    return WMD_ERR_ARG;
  }
  std::vector<int32_t> hq(n + 1), hs(n);
  std::vector<int64_t> ko(n), uo(n);
  int64_t ktot = 0, utot = 0;
  int qmax = 0;
  for (uint64_t Q = 0; Q < n; ++Q) {
    const int64_t q = qp[Q + 1] - qp[Q];
    if (q <= 0 || q > 65536) {
      set_err("query %llu has %lld words", (unsigned long long)Q, (long long)q);
      return WMD_ERR_ARG;
    }
    hq[Q] = (int32_t)qp[Q];
    hs[Q] = round_up4((int)q);
    ko[Q] = ktot;
    uo[Q] = utot;
    ktot += c->Vt * hs[Q];
    utot += c->N * hs[Q];
    qmax = std::max<int>(qmax, (int)q);
  }
  const int64_t total = qp[n];
  if (total >= (1ll << 31)) {
    set_err("too many query words");
    return WMD_ERR_UNSUPPORTED;
  }
  hq[n] = (int32_t)total;
  std::vector<int32_t> words(total);
  std::vector<float> weights(total);
  for (int64_t k = 0; k < total; ++k) {
    if (qw[k] < 0 || qw[k] >= c->V) {
      set_err("query word %lld out of range", (long long)qw[k]);
      return WMD_ERR_ARG;
    }
    if (!(qv[k] > 0.0) || !std::isfinite(qv[k])) {
      set_err("query weight %lld must be finite and > 0", (long long)k);
      return WMD_ERR_ARG;
    }
    words[k] = (int32_t)qw[k];
    weights[k] = (float)qv[k];
  }
  // Distinct query words of the batch and each query word's index among them.
  std::vector<int32_t> dwords(words);
  std::sort(dwords.begin(), dwords.end());
  dwords.erase(std::unique(dwords.begin(), dwords.end()), dwords.end());
  std::vector<int32_t> slot(total);
  for (int64_t k = 0; k < total; ++k)
    slot[k] = (int32_t)(std::lower_bound(dwords.begin(), dwords.end(), words[k]) - dwords.begin());
  const int Sd = (int)dwords.size();
  const int Sdp = (Sd + 63) & ~63;
  const int nqd = Sdp / 64;
  std::vector<int32_t> dqp(nqd + 1), dks(nqd, Sdp);
  std::vector<int64_t> dko(nqd);
  for (int i = 0; i <= nqd; ++i) dqp[i] = std::min(64 * i, Sd);
  for (int i = 0; i < nqd; ++i) dko[i] = 64 * (int64_t)i;
  HIP_TRY(c->d_words.upload(dwords.data(), dwords.size(), c->stream));
  HIP_TRY(c->slot.upload(slot.data(), slot.size(), c->stream));
  HIP_TRY(c->dq_ptr.upload(dqp.data(), dqp.size(), c->stream));
  HIP_TRY(c->dk_stride.upload(dks.data(), dks.size(), c->stream));
  HIP_TRY(c->dk_off.upload(dko.data(), dko.size(), c->stream));
  HIP_TRY(c->q_ptr.upload(hq.data(), hq.size(), c->stream));
  HIP_TRY(c->q_words.upload(words.data(), words.size(), c->stream));
  HIP_TRY(c->q_r.upload(weights.data(), weights.size(), c->stream));
  HIP_TRY(c->k_stride.upload(hs.data(), hs.size(), c->stream));
  HIP_TRY(c->k_off.upload(ko.data(), ko.size(), c->stream));
  HIP_TRY(c->u_off.upload(uo.data(), uo.size(), c->stream));
  std::vector<wmd::StreamQuery> qm(n);
  for (uint64_t Q = 0; Q < n; ++Q)
    qm[Q] = wmd::StreamQuery{hq[Q], (int32_t)(qp[Q + 1] - qp[Q]), hs[Q], 0, ko[Q], uo[Q]};
  HIP_TRY(c->qmeta.upload(qm.data(), qm.size(), c->stream));
  HIP_TRY(c->K.ensure((size_t)ktot));
  HIP_TRY(c->mrow.ensure((size_t)n * c->Vt));
  HIP_TRY(c->wmd.ensure((size_t)n * c->N));
  HIP_TRY(c->iters.ensure((size_t)n * c->N));
  HIP_TRY(c->resid.ensure((size_t)n * c->N));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->nq = (int)n;
  c->qmax = qmax;
  c->Sd = Sd;
  c->Sdp = Sdp;
  c->nqd = nqd;
  c->h_q_ptr = hq;
  c->h_k_stride = hs;
  c->h_k_off = ko;
  c->h_u_off = uo;
  c->k_total = ktot;
  c->u_total = utot;
  c->has_queries = true;
  ++c->plan_epoch;
  c->duals_valid = false;
  c->warm_given = false;
  c->ran = false;
  return WMD_OK;
}

int wmd_context_run(wmd_context* c, double lamda, const wmd_options* opt_in, wmd_stats* stats) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if (!c->has_corpus || !c->has_queries) {
    set_err("corpus and queries must be set before run");
    return WMD_ERR_STATE;
  }
  wmd_options o;
  if (opt_in) o = *opt_in; else wmd_options_default(&o);
  if (!(lamda < 0.0) || !std::isfinite(lamda)) {
    set_err("lamda must be finite and < 0 (K = exp(lamda * M))");
    return WMD_ERR_ARG;
  }
  if (o.max_iter < 0 || o.max_iter > (1 << 24) || !(o.tol >= 0.0)) {
    set_err("bad max_iter/tol");
    return WMD_ERR_ARG;
  }
  if (!(o.omega >= 1.0 && o.omega < 2.0)) {
    set_err("omega must be in [1, 2)");
    return WMD_ERR_ARG;
  }
  if (o.accel != WMD_ACCEL_NONE && o.accel != WMD_ACCEL_ANDERSON) {
    set_err("unknown accel %d", o.accel);
    return WMD_ERR_ARG;
  }
  int variant = o.variant;
  if (variant < WMD_VARIANT_RESIDENT || variant > WMD_VARIANT_NONFUSED) {
    set_err("unknown variant %d", variant);
    return WMD_ERR_ARG;
  }
  if (o.shape >= wmd::kNumShapes) {
    set_err("shape %d out of range", o.shape);
    return WMD_ERR_ARG;
  }
  if (o.domain != WMD_DOMAIN_LINEAR && o.domain != WMD_DOMAIN_LOG) {
    set_err("unknown domain %d", o.domain);
    return WMD_ERR_ARG;
  }
  const bool log_domain = o.domain == WMD_DOMAIN_LOG;
  if (o.cdist != WMD_CDIST_FP32 && o.cdist != WMD_CDIST_MFMA) {
    set_err("unknown cdist mode %d", o.cdist);
    return WMD_ERR_ARG;
  }
  const bool mfma = o.cdist == WMD_CDIST_MFMA;
  if (o.share != WMD_SHARE_AUTO && o.share != WMD_SHARE_OFF && o.share != WMD_SHARE_ON) {
    set_err("unknown share mode %d", o.share);
    return WMD_ERR_ARG;
  }
  const int sq_total = c->h_q_ptr[c->nq];
  const bool share_ok = c->qmax <= wmd::kShareMaxQ &&
                        (size_t)c->Sdp * sizeof(float) <= (size_t)wmd::kShareLds;
  if (o.share == WMD_SHARE_ON && !share_ok) {
    set_err("shared distances need queries of at most %d words and at most %d distinct words "
            "(longest query %d, %d distinct)", wmd::kShareMaxQ,
            wmd::kShareLds / (int)sizeof(float), c->qmax, c->Sd);
    return WMD_ERR_UNSUPPORTED;
  }
  const bool share = o.share == WMD_SHARE_ON ||
                     (o.share == WMD_SHARE_AUTO && share_ok && c->nq > 1 &&
                      5ll * c->Sd <= 4ll * sq_total);
  if (o.warm_start != WMD_WARM_UNIFORM && o.warm_start != WMD_WARM_LEARNED &&
      o.warm_start != WMD_WARM_GIVEN) {
    set_err("unknown warm_start %d", o.warm_start);
    return WMD_ERR_ARG;
  }
  const bool learned = o.warm_start == WMD_WARM_LEARNED;
  const bool given = o.warm_start == WMD_WARM_GIVEN;
  if (given && !c->warm_given) {
    set_err("warm_start = given needs wmd_context_set_warm on the current queries");
    return WMD_ERR_STATE;
  }
  if (learned && !c->mlp_ready) {
    set_err("learned warm start needs MLP parameters (wmd_context_mlp_set / train)");
    return WMD_ERR_STATE;
  }
  if (learned && c->qmax > wmd::kWarmQueryMax) {
    set_err("the learned warm start takes queries of at most %d words (longest %d)",
            wmd::kWarmQueryMax, c->qmax);
    return WMD_ERR_UNSUPPORTED;
  }
  const int Dp = ((int)c->D + 15) & ~15;
  if (mfma && !c->split_ready) {
    // Once per corpus, outside the timed stages: BF16 planes and FP64 norms
    // of every touched word.
    HIP_TRY(c->Es.ensure((size_t)3 * c->Vt * Dp));
    HIP_TRY(c->nt.ensure((size_t)c->Vt));
    wmd::SplitArgs sa{};
    sa.E = c->E.p;
    sa.ids = c->touched.p;
    sa.rows = (int)c->Vt;
    sa.D = (int)c->D;
    sa.Dp = Dp;
    sa.out = c->Es.p;
    sa.norms = c->nt.p;
    HIP_TRY(wmd::launch_split_rows(sa, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->split_ready = true;
  }
  if (learned) {
    rc = warm_corpus_features(c);  // once per corpus, outside the timed stages
    if (rc != WMD_OK) return rc;
  }
  if (variant == WMD_VARIANT_RESIDENT) {
    // Documents beyond the tiles' length run on the streaming kernel (below);
    // a query beyond every tile sends the whole run there.
    bool fits = resident_blocks_fit(c);
    for (int Q = 0; Q < c->nq && fits; ++Q) {
      const int q = c->h_q_ptr[Q + 1] - c->h_q_ptr[Q];
      fits = o.shape >= 0 ? shape_fits(o.shape, q, log_domain) : pick_shape(q, 0, log_domain) >= 0;
    }
    if (!fits) {
      if (o.shape >= 0) {
        set_err(resident_blocks_fit(c)
                    ? "tile shape %d cannot hold the longest query (%d words), or has no "
                      "log-domain form"
                    : "tile shape %d: a query's kernel block (%d-word query) exceeds the resident "
                      "solver's 2 GiB block or 2^24-word limit",
                o.shape, c->qmax);
        return WMD_ERR_UNSUPPORTED;
      }
      variant = WMD_VARIANT_FUSED;  // queries beyond the register tiles or blocks
    }
  }
  const bool long_docs = variant == WMD_VARIANT_RESIDENT && c->n_long > 0;
  if (long_docs && wmd::long_lds_bytes(round_up4(c->qmax), round_up4(c->nmax)) > 64 * 1024) {
    set_err("documents of %d words are too long for the long-document kernel", c->nmax);
    return WMD_ERR_UNSUPPORTED;
  }
  if ((learned || given || o.export_duals) && (variant != WMD_VARIANT_RESIDENT || log_domain)) {
    set_err("learned warm start and dual export run on the resident variant in the linear domain");
    return WMD_ERR_UNSUPPORTED;
  }
  if (o.accel != WMD_ACCEL_NONE && (variant != WMD_VARIANT_RESIDENT || !(o.tol > 0.0))) {
    set_err("Anderson acceleration runs on the resident variant with a tolerance (tol > 0)");
    return WMD_ERR_UNSUPPORTED;
  }
  if (log_domain && variant == WMD_VARIANT_NONFUSED) {
    set_err("the log domain runs on the resident and fused variants");
    return WMD_ERR_UNSUPPORTED;
  }
  if (variant != WMD_VARIANT_RESIDENT) {
    const size_t lds = wmd::stream_lds_bytes(round_up4(std::max(c->qmax, 1)),
                                             round_up4(std::max(c->nmax, 1)));
    if (lds > 64 * 1024) {
      set_err("query (%d words) or document (%d words) too long for the streaming kernels",
              c->qmax, c->nmax);
      return WMD_ERR_UNSUPPORTED;
    }
  }

  // Shared form: the distance kernels run in raw mode over pseudo-queries of
  // 64 distinct words into Dd, then k_share_rescale builds every query's block.
  const int32_t* cd_words = share ? c->d_words.p : c->q_words.p;
  const int32_t* cd_ptr = share ? c->dq_ptr.p : c->q_ptr.p;
  const int64_t* cd_off = share ? c->dk_off.p : c->k_off.p;
  const int32_t* cd_stride = share ? c->dk_stride.p : c->k_stride.p;
  const int cd_nq = share ? c->nqd : c->nq;
  const int cd_rows = share ? c->Sd : sq_total;
  if (share) HIP_TRY(c->Dd.ensure((size_t)c->Vt * (size_t)c->Sdp));
  float* cd_out = share ? c->Dd.p : c->K.p;

  wmd::CdistArgs ca{};
  ca.E = c->E.p;
  ca.touched = c->touched.p;
  ca.q_ptr = cd_ptr;
  ca.q_words = cd_words;
  ca.k_off = cd_off;
  ca.k_stride = cd_stride;
  ca.K = cd_out;
  ca.mrow = c->mrow.p;
  ca.D = (int)c->D;
  ca.Vt = (int)c->Vt;
  ca.lamda = (float)lamda;
  ca.log_out = log_domain ? 1 : 0;
  ca.raw = share ? 1 : 0;

  int launches = 0;
  HIP_TRY(hipEventRecord(c->ev[0], c->stream));
  if (mfma) {
    const int sq = cd_rows;
    HIP_TRY(c->Qs.ensure((size_t)3 * sq * Dp));
    HIP_TRY(c->nqw.ensure((size_t)sq));
    wmd::SplitArgs sa{};
    sa.E = c->E.p;
    sa.ids = cd_words;
    sa.rows = sq;
    sa.D = (int)c->D;
    sa.Dp = Dp;
    sa.out = c->Qs.p;
    sa.norms = c->nqw.p;
    HIP_TRY(wmd::launch_split_rows(sa, c->stream));
    wmd::MfmaCdistArgs ma{};
    ma.Es = c->Es.p;
    ma.Qs = c->Qs.p;
    ma.nt = c->nt.p;
    ma.nq = c->nqw.p;
    ma.E = c->E.p;
    ma.touched = c->touched.p;
    ma.q_words = cd_words;
    ma.q_ptr = cd_ptr;
    ma.k_off = cd_off;
    ma.k_stride = cd_stride;
    ma.K = cd_out;
    ma.mrow = c->mrow.p;
    ma.D = (int)c->D;
    ma.Dp = Dp;
    ma.Vt = (int)c->Vt;
    ma.Sq = sq;
    ma.lamda = (float)lamda;
    ma.log_out = log_domain ? 1 : 0;
    ma.raw = share ? 1 : 0;
    HIP_TRY(wmd::launch_cdist_mfma(ma, cd_nq, c->stream));
    launches += 2;
  } else {
    HIP_TRY(wmd::launch_cdist(ca, cd_nq, c->stream));
    ++launches;
  }
  if (share) {
    wmd::ShareArgs sh{};
    sh.Dd = c->Dd.p;
    sh.slot = c->slot.p;
    sh.q_ptr = c->q_ptr.p;
    sh.k_off = c->k_off.p;
    sh.k_stride = c->k_stride.p;
    sh.K = c->K.p;
    sh.mrow = c->mrow.p;
    sh.Vt = (int)c->Vt;
    sh.Sdp = c->Sdp;
    sh.nq = c->nq;
    // Rows per block: as many as fit in 20 KB of LDS (eight blocks per CU),
    // at least one.
    sh.rows = std::max(1, std::min(16, 20 * 1024 / (c->Sdp * (int)sizeof(float))));
    sh.lamda = (float)lamda;
    sh.log_out = log_domain ? 1 : 0;
    sh.sum_q = c->h_q_ptr[c->nq];
    HIP_TRY(wmd::launch_share_rescale(sh, c->stream));
    ++launches;
  }
  HIP_TRY(hipEventRecord(c->ev[1], c->stream));
  wmd::SolveArgs a = base_args(c, lamda, o);
  if (learned) {
    rc = mlp_predict(c, lamda, &launches);
    if (rc != WMD_OK) return rc;
    a.warm = c->pred.p;
  } else if (given) {
    a.warm = c->warm_vals.p;
  }
  HIP_TRY(c->fallbacks.ensure(1));
  HIP_TRY(wmd::launch_zero_u32(c->fallbacks.p, 1, c->stream));
  a.fallbacks = c->fallbacks.p;
  if (o.export_duals) {
    HIP_TRY(c->duals.ensure((size_t)c->nq * (size_t)c->nnz));
    a.duals = c->duals.p;
  }
  a.log_domain = log_domain ? 1 : 0;
  HIP_TRY(hipEventRecord(c->ev[3], c->stream));
  if (variant == WMD_VARIANT_RESIDENT)
    rc = run_resident(c, a, o.shape, log_domain, o.graph != 0, &launches);
  else
    rc = run_streaming(c, a, variant == WMD_VARIANT_FUSED, o, &launches);
  if (rc != WMD_OK) return rc;
  HIP_TRY(hipEventRecord(c->ev[2], c->stream));
  HIP_TRY(hipEventSynchronize(c->ev[2]));
  c->ran = true;
  c->last_lamda = lamda;
  c->duals_valid = o.export_duals != 0;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    float t_all = 0.f, t_cd = 0.f, t_sv = 0.f, t_w = 0.f;
    HIP_TRY(hipEventElapsedTime(&t_all, c->ev[0], c->ev[2]));
    HIP_TRY(hipEventElapsedTime(&t_cd, c->ev[0], c->ev[1]));
    HIP_TRY(hipEventElapsedTime(&t_w, c->ev[1], c->ev[3]));
    HIP_TRY(hipEventElapsedTime(&t_sv, c->ev[3], c->ev[2]));
    stats->ms_total = t_all;
    stats->ms_cdist = t_cd;
    stats->ms_warm = t_w;
    stats->ms_solve = t_sv;
    uint32_t fb = 0;
    HIP_TRY(hipMemcpy(&fb, c->fallbacks.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
    stats->fallbacks = (int32_t)fb;
    stats->pairs = (int64_t)c->nq * c->N;
    stats->launches = launches;
    stats->variant = variant;
    stats->shared = share ? 1 : 0;
    stats->distinct = c->Sd;
    stats->streamed_docs = variant == WMD_VARIANT_RESIDENT ? (long_docs ? c->n_long : 0)
                                                           : (int32_t)c->N;
    if (!o.summarize) return WMD_OK;
    const size_t np = (size_t)c->nq * c->N;
    std::vector<int32_t> it(np);
    std::vector<float> rs(np);
    HIP_TRY(hipMemcpy(it.data(), c->iters.p, np * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(rs.data(), c->resid.p, np * sizeof(float), hipMemcpyDeviceToHost));
    if (variant == WMD_VARIANT_RESIDENT) measure_plan(c, it);
    double sum = 0.0, rmax = 0.0;
    int imax = 0;
    for (size_t k = 0; k < np; ++k) {
      sum += it[k];
      imax = std::max(imax, it[k]);
      if (std::isnan(rs[k]) || rs[k] > rmax) rmax = rs[k];
    }
    stats->iters_max = imax;
    stats->iters_mean = np ? sum / (double)np : 0.0;
    stats->resid_max = rmax;
  }
  return WMD_OK;
}

int wmd_tuning_dims(int32_t* n_classes, int32_t* max_q) {
  if (n_classes) *n_classes = wmd::kNumDocClasses;
  if (max_q) *max_q = wmd::kMaxQ;
  return WMD_OK;
}

namespace {

// A tuning table is [kNumDocClasses][kMaxQ + 1]; a caller's buffer of any
// other length (e.g. saved by a build with other class or shape limits) is
// refused before it is read or written.
int check_tuning_len(uint64_t n_entries) {
  const uint64_t want = (uint64_t)wmd::kNumDocClasses * kTuneStride;
  if (n_entries != want) {
    set_err("tuning table has %llu entries, this build's is %d classes x %d = %llu "
            "(wmd_tuning_dims)", (unsigned long long)n_entries, wmd::kNumDocClasses,
            kTuneStride, (unsigned long long)want);
    return WMD_ERR_ARG;
  }
  return WMD_OK;
}

std::vector<int32_t> default_tuning(bool log_domain) {
  std::vector<int32_t> t((size_t)wmd::kNumDocClasses * kTuneStride, -1);
  for (int cls = 0; cls < wmd::kNumDocClasses; ++cls)
    for (int q = 1; q <= wmd::kMaxQ; ++q) t[(size_t)cls * kTuneStride + q] = pick_shape(q, 0, log_domain);
  return t;
}

}  // namespace

int wmd_context_autotune(wmd_context* c, double lamda, const wmd_options* opt_in, int32_t iters,
                         double* ms_out) {
  const auto t_start = std::chrono::steady_clock::now();
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  wmd_options o;
  if (opt_in) o = *opt_in; else wmd_options_default(&o);
  if (iters < 1 || iters > 4096) {
    set_err("tuning iterations must be in [1, 4096]");
    return WMD_ERR_ARG;
  }
  // The distance stage of a zero-iteration run leaves K and mrow for lamda in
  // the options' domain (and validates corpus, queries and options).  The
  // candidates are then timed with the options' update rule (plain,
  // over-relaxed or Anderson), whose per-iteration cost differs.
  const int accel = o.accel;
  const double omega = o.omega;
  o.variant = WMD_VARIANT_RESIDENT;
  o.max_iter = 0;
  o.tol = 0.0;
  o.accel = WMD_ACCEL_NONE;
  o.shape = -1;
  o.warm_start = WMD_WARM_UNIFORM;
  o.export_duals = 0;
  o.summarize = 0;
  if (c->has_corpus && c->has_queries && c->qmax > wmd::kMaxQ) {
    set_err("queries (%d words) beyond the resident tiles", c->qmax);
    return WMD_ERR_UNSUPPORTED;
  }
  rc = wmd_context_run(c, lamda, &o, nullptr);
  if (rc != WMD_OK) return rc;
  const bool lg = o.domain == WMD_DOMAIN_LOG;
  o.max_iter = iters;
  o.omega = omega;
  o.accel = accel;
  wmd::SolveArgs base = base_args(c, lamda, o);
  base.fallbacks = c->fallbacks.p;
  // A tolerance no pair reaches keeps every candidate at `iters` iterations
  // (the accelerated rule needs one to run).
  if (accel != WMD_ACCEL_NONE) base.tol = 1e-30f;

  std::vector<int32_t> table = default_tuning(lg);
  std::vector<int32_t> ids;
  for (int cls = 0; cls < wmd::kNumDocClasses; ++cls) {
    const int64_t d0 = c->h_class_off[cls], d1 = c->h_class_off[cls + 1];
    if (d1 == d0) continue;
    // Buckets: queries sharing their smallest-fitting shape.
    std::vector<std::vector<int32_t>> groups(wmd::kNumShapes);
    for (int Q = 0; Q < c->nq; ++Q) {
      const int q = c->h_q_ptr[Q + 1] - c->h_q_ptr[Q];
      groups[pick_shape(q, 0, lg)].push_back(Q);
    }
    int prev_cap = 0;
    for (int s0 = 0; s0 < wmd::kNumShapes; ++s0) {
      if (!shape_in_domain(s0, lg)) continue;
      const int cap0 = wmd::kShapes[s0].cap_q();
      const std::vector<int32_t>& g = groups[s0];
      if (!g.empty()) {
        int qmax = 0;
        for (int Q : g) qmax = std::max(qmax, c->h_q_ptr[Q + 1] - c->h_q_ptr[Q]);
        // Sample of the class's documents: enough pairs to fill the device.
        const int64_t n_docs = std::min<int64_t>(
            d1 - d0, std::max<int64_t>(128, 65536 / (int64_t)g.size()));
        HIP_TRY(c->bucket.upload(g.data(), g.size(), c->stream));
        ResidentLaunch l{0, cls, 0, (int64_t)g.size(), d0, d0 + n_docs, 0.0};
        float best_ms = INFINITY;
        int best = s0, tried = 0;
        for (int s = s0; s < wmd::kNumShapes && tried < kTuneCandidates; ++s) {
          if (!shape_fits(s, qmax, lg) || wmd::kShapes[s].cap_q() > 2 * cap0) continue;
          ++tried;
          l.shape = s;
          HIP_TRY(launch_one(c, base, l, c->bucket.p, lg, c->stream));  // code load, caches
          float ms = INFINITY;
          for (int rep = 0; rep < kTuneReps; ++rep) {
            HIP_TRY(hipEventRecord(c->ev[0], c->stream));
            HIP_TRY(launch_one(c, base, l, c->bucket.p, lg, c->stream));
            HIP_TRY(hipEventRecord(c->ev[1], c->stream));
            HIP_TRY(hipEventSynchronize(c->ev[1]));
            float t = 0.f;
            HIP_TRY(hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
            ms = std::min(ms, t);
          }
          // Another shape replaces the default only when clearly faster: a
          // launch timed alone differs from the same launch beside the
          // plan's others, and small differences are noise.
          if (s == s0 ? ms < best_ms : ms < kTuneMargin * best_ms) {
            best_ms = ms;
            best = s;
          }
        }
        for (int q = prev_cap + 1; q <= cap0; ++q) table[(size_t)cls * kTuneStride + q] = best;
      }
      prev_cap = cap0;
    }
  }
  c->tune[lg ? 1 : 0] = table;
  ++c->plan_epoch;
  c->ran = false;  // the tuning launches overwrote the results
  if (ms_out)
    *ms_out = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  return WMD_OK;
}

int wmd_context_get_tuning(wmd_context* c, int32_t domain, int32_t* table, uint64_t n_entries) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if (!table || (domain != WMD_DOMAIN_LINEAR && domain != WMD_DOMAIN_LOG)) {
    set_err("null table or unknown domain");
    return WMD_ERR_ARG;
  }
  rc = check_tuning_len(n_entries);
  if (rc != WMD_OK) return rc;
  const bool lg = domain == WMD_DOMAIN_LOG;
  const std::vector<int32_t> t = c->tune[lg ? 1 : 0].empty() ? default_tuning(lg) : c->tune[lg ? 1 : 0];
  std::memcpy(table, t.data(), t.size() * sizeof(int32_t));
  return WMD_OK;
}

int wmd_context_set_tuning(wmd_context* c, int32_t domain, const int32_t* table,
                           uint64_t n_entries) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if (domain != WMD_DOMAIN_LINEAR && domain != WMD_DOMAIN_LOG) {
    set_err("unknown domain %d", domain);
    return WMD_ERR_ARG;
  }
  const bool lg = domain == WMD_DOMAIN_LOG;
  if (!table) {
    c->tune[lg ? 1 : 0].clear();
    ++c->plan_epoch;
    return WMD_OK;
  }
  rc = check_tuning_len(n_entries);
  if (rc != WMD_OK) return rc;
  for (int cls = 0; cls < wmd::kNumDocClasses; ++cls)
    for (int q = 1; q <= wmd::kMaxQ; ++q) {
      const int s = table[(size_t)cls * kTuneStride + q];
      if (s != -1 && !shape_fits(s, q, lg)) {
        set_err("tuning entry (class %d, q %d) = %d does not hold q words%s", cls, q, s,
                lg ? " or has no log-domain form" : "");
        return WMD_ERR_ARG;
      }
    }
  c->tune[lg ? 1 : 0].assign(table, table + (size_t)wmd::kNumDocClasses * kTuneStride);
  ++c->plan_epoch;
  return WMD_OK;
}

int wmd_context_run_query_noalloc(wmd_context* c, int32_t q, const int32_t* q_words,
                                  const float* q_weights, double lamda, const wmd_options* opt_in,
                                  double* wmd_out, int32_t* iters_out, float* resid_out) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if (!c->has_corpus) {
    set_err("set_corpus must be called before a query");
    return WMD_ERR_STATE;
  }
  if (!q_words || !q_weights || !wmd_out || q < 1 || q > wmd::kMaxQ) {
    set_err("need device arrays of 1..%d query words and weights and a result array (q=%d)",
            wmd::kMaxQ, q);
    return WMD_ERR_ARG;
  }
  wmd_options o;
  if (opt_in) o = *opt_in; else wmd_options_default(&o);
  if (!(lamda < 0.0) || !std::isfinite(lamda)) {
    set_err("lamda must be finite and < 0 (K = exp(lamda * M))");
    return WMD_ERR_ARG;
  }
  if (o.max_iter < 0 || o.max_iter > (1 << 24) || !(o.tol >= 0.0) ||
      !(o.omega >= 1.0 && o.omega < 2.0) ||
      (o.accel != WMD_ACCEL_NONE && o.accel != WMD_ACCEL_ANDERSON) ||
      (o.domain != WMD_DOMAIN_LINEAR && o.domain != WMD_DOMAIN_LOG) ||
      (o.cdist != WMD_CDIST_FP32 && o.cdist != WMD_CDIST_MFMA) || o.shape >= wmd::kNumShapes) {
    set_err("bad options");
    return WMD_ERR_ARG;
  }
  if (o.variant != WMD_VARIANT_RESIDENT || o.warm_start != WMD_WARM_UNIFORM || o.export_duals) {
    set_err("the allocation-free entry runs the resident variant from the uniform start");
    return WMD_ERR_UNSUPPORTED;
  }
  if (o.accel != WMD_ACCEL_NONE && !(o.tol > 0.0)) {
    set_err("Anderson acceleration runs with a tolerance (tol > 0)");
    return WMD_ERR_UNSUPPORTED;
  }
  const bool log_domain = o.domain == WMD_DOMAIN_LOG;
  const bool mfma = o.cdist == WMD_CDIST_MFMA;
  if (c->nmax > wmd::kShapes[0].cap_n()) {
    set_err("documents of %d words are beyond the resident tiles (%d)", c->nmax,
            wmd::kShapes[0].cap_n());
    return WMD_ERR_UNSUPPORTED;
  }
  if (o.shape >= 0 ? !shape_fits(o.shape, q, log_domain) : pick_shape(q, 0, log_domain) < 0) {
    set_err("no tile shape of the domain holds %d query words", q);
    return WMD_ERR_UNSUPPORTED;
  }
  if (mfma && !c->split_ready) {
    set_err("the matrix-core distance stage needs the corpus planes of an earlier "
            "wmd_context_run with cdist = WMD_CDIST_MFMA (or use WMD_CDIST_FP32)");
    return WMD_ERR_STATE;
  }
  // Scratch layout, every part 256-byte aligned.
  const int ks = round_up4(q);
  const int Dp = ((int)c->D + 15) & ~15;
  size_t at = 0;
  auto take = [&](size_t bytes) {
    const size_t off = at;
    at += (bytes + 255) & ~(size_t)255;
    return off;
  };
  const size_t o_words = take((size_t)q * 4), o_weights = take((size_t)q * 4);
  const size_t o_qptr = take(8), o_koff = take(8), o_kstride = take(4), o_bucket = take(4);
  const size_t o_fb = take(4), o_status = take(4);
  const size_t o_K = take((size_t)c->Vt * ks * 4), o_mrow = take((size_t)c->Vt * 4);
  const size_t o_Qs = mfma ? take((size_t)3 * q * Dp * 2) : 0;
  const size_t o_nqw = mfma ? take((size_t)q * 8) : 0;
  const size_t o_iters = iters_out ? 0 : take((size_t)c->N * 4);
  const size_t o_resid = resid_out ? 0 : take((size_t)c->N * 4);
  if (at > kScratchBytes) {
    set_err("this query needs %zu bytes of scratch (kernel block %lld x %d), the scratch holds "
            "%zu", at, (long long)c->Vt, ks, kScratchBytes);
    return WMD_ERR_UNSUPPORTED;
  }
  unsigned char* S = c->scratch;
  int32_t* words = reinterpret_cast<int32_t*>(S + o_words);
  float* weights = reinterpret_cast<float*>(S + o_weights);
  int32_t* qptr = reinterpret_cast<int32_t*>(S + o_qptr);
  int64_t* koff = reinterpret_cast<int64_t*>(S + o_koff);
  int32_t* kstride = reinterpret_cast<int32_t*>(S + o_kstride);
  int32_t* bucket = reinterpret_cast<int32_t*>(S + o_bucket);
  uint32_t* fb = reinterpret_cast<uint32_t*>(S + o_fb);
  uint32_t* status = reinterpret_cast<uint32_t*>(S + o_status);
  float* K = reinterpret_cast<float*>(S + o_K);
  float* mrow = reinterpret_cast<float*>(S + o_mrow);

  wmd::NoAllocSetupArgs sa{};
  sa.words_in = q_words;
  sa.weights_in = q_weights;
  sa.q = q;
  sa.V = c->V;
  sa.words = words;
  sa.weights = weights;
  sa.q_ptr = qptr;
  sa.k_off = koff;
  sa.k_stride = kstride;
  sa.bucket = bucket;
  sa.fallbacks = fb;
  sa.status = status;
  HIP_TRY(wmd::launch_noalloc_setup(sa, c->stream));
  // Distance stage for the one query: its words are the only rows.
  if (mfma) {
    uint16_t* Qs = reinterpret_cast<uint16_t*>(S + o_Qs);
    double* nqw = reinterpret_cast<double*>(S + o_nqw);
    wmd::SplitArgs sp{};
    sp.E = c->E.p;
    sp.ids = words;
    sp.rows = q;
    sp.D = (int)c->D;
    sp.Dp = Dp;
    sp.out = Qs;
    sp.norms = nqw;
    HIP_TRY(wmd::launch_split_rows(sp, c->stream));
    wmd::MfmaCdistArgs ma{};
    ma.Es = c->Es.p;
    ma.Qs = Qs;
    ma.nt = c->nt.p;
    ma.nq = nqw;
    ma.E = c->E.p;
    ma.touched = c->touched.p;
    ma.q_words = words;
    ma.q_ptr = qptr;
    ma.k_off = koff;
    ma.k_stride = kstride;
    ma.K = K;
    ma.mrow = mrow;
    ma.D = (int)c->D;
    ma.Dp = Dp;
    ma.Vt = (int)c->Vt;
    ma.Sq = q;
    ma.lamda = (float)lamda;
    ma.log_out = log_domain ? 1 : 0;
    HIP_TRY(wmd::launch_cdist_mfma(ma, 1, c->stream));
  } else {
    wmd::CdistArgs ca{};
    ca.E = c->E.p;
    ca.touched = c->touched.p;
    ca.q_ptr = qptr;
    ca.q_words = words;
    ca.k_off = koff;
    ca.k_stride = kstride;
    ca.K = K;
    ca.mrow = mrow;
    ca.D = (int)c->D;
    ca.Vt = (int)c->Vt;
    ca.lamda = (float)lamda;
    ca.log_out = log_domain ? 1 : 0;
    HIP_TRY(wmd::launch_cdist(ca, 1, c->stream));
  }
  // Solver: one launch per document-length class, the shapes wmd_context_run
  // picks for this query length.
  wmd::SolveArgs a{};
  a.K = K;
  a.mrow = mrow;
  a.k_off = koff;
  a.k_stride = kstride;
  a.q_ptr = qptr;
  a.q_r = weights;
  a.doc_ptr = c->doc_ptr.p;
  a.doc_t = c->doc_t.p;
  a.doc_t1 = c->doc_t1.p;
  a.doc_val = c->doc_val.p;
  a.nnz = c->nnz;
  a.N = (int)c->N;
  a.Vt = (int)c->Vt;
  a.max_iter = o.max_iter;
  a.tol = (float)o.tol;
  a.omega = (float)o.omega;
  a.accel = o.accel;
  a.inv_lamda = (float)(1.0 / lamda);
  a.lamda = (float)lamda;
  a.wmd = wmd_out;
  a.iters = iters_out ? iters_out : reinterpret_cast<int32_t*>(S + o_iters);
  a.resid = resid_out ? resid_out : reinterpret_cast<float*>(S + o_resid);
  a.fallbacks = fb;
  a.lds_q = ks;
  a.lds_n = round_up4(std::max(c->nmax, 1));
  for (int cls = 0; cls < wmd::kNumDocClasses; ++cls) {
    const int64_t d0 = c->h_class_off[cls], d1 = c->h_class_off[cls + 1];
    if (d1 == d0) continue;
    ResidentLaunch l{shape_for(c, cls, q, log_domain, o.shape), cls, 0, 1, d0, d1, 0.0};
    HIP_TRY(launch_one(c, a, l, bucket, log_domain, c->stream));
  }
  HIP_TRY(hipMemcpyAsync(c->h_status, status, sizeof(uint32_t), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->h_status[0] != 0) {
    set_err("a query word is outside [0, %lld) or a weight is not finite and > 0; the results "
            "are not valid", (long long)c->V);
    return WMD_ERR_ARG;
  }
  return WMD_OK;
}

int wmd_context_result(wmd_context* c, double* out, int32_t* iters, float* resid) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if (!c->ran) {
    set_err("run must complete before result");
    return WMD_ERR_STATE;
  }
  if (!out) {
    set_err("null output");
    return WMD_ERR_ARG;
  }
  const size_t np = (size_t)c->nq * c->N;
  HIP_TRY(hipMemcpy(out, c->wmd.p, np * sizeof(double), hipMemcpyDeviceToHost));
  if (iters) HIP_TRY(hipMemcpy(iters, c->iters.p, np * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (resid) HIP_TRY(hipMemcpy(resid, c->resid.p, np * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t k = 0; k < np; ++k)
    if (!std::isfinite(out[k])) {
      set_err("result %zu (query %zu, doc %zu) is not finite", k, k / (size_t)c->N,
              k % (size_t)c->N);
      return WMD_ERR_NONFINITE;
    }
  return WMD_OK;
}

const double* wmd_context_device_result(const wmd_context* c) {
  return (c && c->ran) ? c->wmd.p : nullptr;
}

int wmd_context_topk(wmd_context* c, int32_t k, int32_t* doc_ids, double* wmd_out) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if (!c->ran) {
    set_err("run must complete before topk");
    return WMD_ERR_STATE;
  }
  if (k < 1 || k > c->N || !doc_ids || !wmd_out) {
    set_err("k must be in [1, N = %lld] and outputs non-null", (long long)c->N);
    return WMD_ERR_ARG;
  }
  const size_t total = (size_t)c->nq * c->N, nk = (size_t)c->nq * k;
  if (total > (size_t)INT32_MAX) {
    set_err("queries x documents beyond 2^31 for the segmented sort");
    return WMD_ERR_UNSUPPORTED;
  }
  DevBuf<double> keys_in, keys_out, dist;
  DevBuf<int32_t> vals_in, vals_out, offsets, idx;
  DevBuf<uint8_t> temp;
  wmd::TopkArgs t{};
  t.temp_bytes = wmd::topk_temp_bytes(c->nq, (int)c->N);
  HIP_TRY(keys_in.ensure(total));
  HIP_TRY(keys_out.ensure(total));
  HIP_TRY(vals_in.ensure(total));
  HIP_TRY(vals_out.ensure(total));
  HIP_TRY(offsets.ensure((size_t)c->nq + 1));
  HIP_TRY(temp.ensure(t.temp_bytes));
  HIP_TRY(dist.ensure(nk));
  HIP_TRY(idx.ensure(nk));
  t.wmd = c->wmd.p;
  t.nq = c->nq;
  t.N = (int)c->N;
  t.k = k;
  t.keys_in = keys_in.p;
  t.keys_out = keys_out.p;
  t.vals_in = vals_in.p;
  t.vals_out = vals_out.p;
  t.offsets = offsets.p;
  t.temp = temp.p;
  t.dist = dist.p;
  t.idx = idx.p;
  hipError_t e = wmd::launch_topk(t, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(doc_ids, idx.p, nk * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(wmd_out, dist.p, nk * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  keys_in.release(); keys_out.release(); vals_in.release(); vals_out.release();
  offsets.release(); temp.release(); dist.release(); idx.release();
  if (e != hipSuccess) {
    set_err("topk failed: %s", hipGetErrorString(e));
    return WMD_ERR_HIP;
  }
  return WMD_OK;
}

int wmd_sinkhorn_ex(const double* r, const double* vecs, double* WMD, double lamda, uint64_t V,
                    uint64_t N, uint64_t D, const uint64_t* row_ptr, const uint64_t* col_inds,
                    const double* vals, const wmd_options* opt, wmd_stats* stats) {
  return one_shot(r, vecs, WMD, lamda, V, N, D, row_ptr, col_inds, vals, opt, stats);
}

int wmd_sinkhorn(const double* r, const double* vecs, double* WMD, double lamda, int max_iter,
                 uint64_t V, uint64_t N, uint64_t D, const uint64_t* row_ptr,
                 const uint64_t* col_inds, const double* vals) {
  wmd_options o;
  wmd_options_default(&o);
  o.max_iter = max_iter;
  return one_shot(r, vecs, WMD, lamda, V, N, D, row_ptr, col_inds, vals, &o, nullptr);
}

int wmd_sinkhorn_nonfused(const double* r, const double* vecs, double* WMD, double lamda,
                          int max_iter, uint64_t V, uint64_t N, uint64_t D,
                          const uint64_t* row_ptr, const uint64_t* col_inds, const double* vals) {
  if (!(lamda > 0.0)) {
    set_err("nonfused convention K = exp(-M * lamda) needs lamda > 0");
    return WMD_ERR_ARG;
  }
  wmd_options o;
  wmd_options_default(&o);
  o.max_iter = max_iter;
  o.variant = WMD_VARIANT_NONFUSED;
  return one_shot(r, vecs, WMD, -lamda, V, N, D, row_ptr, col_inds, vals, &o, nullptr);
}


uint64_t wmd_mlp_param_count(void) { return (uint64_t)wmd::kWarmParams; }

void wmd_train_options_default(wmd_train_options* opt) {
  if (!opt) return;
  std::memset(opt, 0, sizeof(*opt));
  opt->epochs = 30;
  opt->pairs_per_batch = 64;
  opt->lr = 3e-3;
  opt->weight_decay = 0.0;
  opt->seed = 0;
  opt->standardize = 1;
}

int wmd_context_mlp_set(wmd_context* c, const float* params, uint64_t n_params) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if (!params || n_params != (uint64_t)wmd::kWarmParams) {
    set_err("expected %d MLP parameters", wmd::kWarmParams);
    return WMD_ERR_ARG;
  }
  for (uint64_t k = 0; k < n_params; ++k)
    if (!std::isfinite(params[k])) {
      set_err("MLP parameter %llu is not finite", (unsigned long long)k);
      return WMD_ERR_ARG;
    }
  HIP_TRY(c->mlp_p.upload(params, n_params, c->stream));
  HIP_TRY(c->mlp_m.ensure(n_params));
  HIP_TRY(c->mlp_v.ensure(n_params));
  HIP_TRY(hipMemsetAsync(c->mlp_m.p, 0, n_params * sizeof(float), c->stream));
  HIP_TRY(hipMemsetAsync(c->mlp_v.p, 0, n_params * sizeof(float), c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->adam_step = 0;
  c->mlp_ready = true;
  return WMD_OK;
}

int wmd_context_mlp_get(wmd_context* c, float* params, uint64_t n_params) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if (!c->mlp_ready) {
    set_err("no MLP parameters in this context");
    return WMD_ERR_STATE;
  }
  if (!params || n_params != (uint64_t)wmd::kWarmParams) {
    set_err("expected %d MLP parameters", wmd::kWarmParams);
    return WMD_ERR_ARG;
  }
  HIP_TRY(hipMemcpy(params, c->mlp_p.p, n_params * sizeof(float), hipMemcpyDeviceToHost));
  return WMD_OK;
}

int wmd_context_set_warm(wmd_context* c, const float* log_v) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if (!c->has_queries) {
    set_err("set_queries must come first");
    return WMD_ERR_STATE;
  }
  if (!log_v) {
    set_err("null input");
    return WMD_ERR_ARG;
  }
  const size_t n = (size_t)c->nq * (size_t)c->nnz;
  for (size_t k = 0; k < n; ++k)
    if (!std::isfinite(log_v[k])) {
      set_err("warm value %zu is not finite", k);
      return WMD_ERR_ARG;
    }
  HIP_TRY(c->warm_vals.upload(log_v, n, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->warm_given = true;
  return WMD_OK;
}

int wmd_context_duals(wmd_context* c, float* out) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if (!c->duals_valid) {
    set_err("no exported duals: run with export_duals = 1 on the current corpus and queries");
    return WMD_ERR_STATE;
  }
  if (!out) {
    set_err("null output");
    return WMD_ERR_ARG;
  }
  HIP_TRY(hipMemcpy(out, c->duals.p, (size_t)c->nq * c->nnz * sizeof(float),
                    hipMemcpyDeviceToHost));
  return WMD_OK;
}

int wmd_context_warm_features(wmd_context* c, float* out) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if (!c->ran) {
    set_err("run must complete before the warm-start features (they use its row minima)");
    return WMD_ERR_STATE;
  }
  if (!out) {
    set_err("null output");
    return WMD_ERR_ARG;
  }
  rc = warm_corpus_features(c);
  if (rc != WMD_OK) return rc;
  int launches = 0;
  rc = warm_query_features(c, &launches);
  if (rc != WMD_OK) return rc;
  const size_t n = (size_t)c->nq * (size_t)c->nnz * wmd::kWarmFeatures;
  DevBuf<float> feats;
  HIP_TRY(feats.ensure(n));
  wmd::WarmInferArgs a = warm_args(c, c->last_lamda);
  a.feats = feats.p;
  hipError_t e = wmd::launch_warm_infer(a, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, feats.p, n * sizeof(float), hipMemcpyDeviceToHost,
                                          c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  feats.release();
  if (e != hipSuccess) {
    set_err("warm-start features failed: %s", hipGetErrorString(e));
    return WMD_ERR_HIP;
  }
  return WMD_OK;
}

int wmd_context_warm_predictions(wmd_context* c, float* out) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if (!c->ran) {
    set_err("run must complete before the warm-start predictions (they use its row minima)");
    return WMD_ERR_STATE;
  }
  if (!c->mlp_ready) {
    set_err("set MLP parameters first (wmd_context_mlp_set)");
    return WMD_ERR_STATE;
  }
  if (!out) {
    set_err("null output");
    return WMD_ERR_ARG;
  }
  rc = warm_corpus_features(c);
  if (rc != WMD_OK) return rc;
  int launches = 0;
  rc = mlp_predict(c, c->last_lamda, &launches);
  if (rc != WMD_OK) return rc;
  hipError_t e = hipMemcpyAsync(out, c->pred.p, sizeof(float) * (size_t)c->nq * (size_t)c->nnz,
                                hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    set_err("warm-start predictions failed: %s", hipGetErrorString(e));
    return WMD_ERR_HIP;
  }
  return WMD_OK;
}

int wmd_context_mlp_train(wmd_context* c, const wmd_train_options* opt_in, double* epoch_loss) {
  int rc = check_ctx(c);
  if (rc != WMD_OK) return rc;
  if (!c->mlp_ready) {
    set_err("set MLP parameters first (wmd_context_mlp_set)");
    return WMD_ERR_STATE;
  }
  if (!c->duals_valid) {
    set_err("training targets missing: run with export_duals = 1 on the current batch first");
    return WMD_ERR_STATE;
  }
  wmd_train_options o;
  if (opt_in) o = *opt_in; else wmd_train_options_default(&o);
  if (o.epochs < 0 || o.pairs_per_batch <= 0 || o.pairs_per_batch > 4096 || !(o.lr > 0.0) ||
      !(o.weight_decay >= 0.0)) {
    set_err("bad training options");
    return WMD_ERR_ARG;
  }
  rc = warm_corpus_features(c);
  if (rc != WMD_OK) return rc;
  int launches = 0;
  rc = warm_query_features(c, &launches);
  if (rc != WMD_OK) return rc;
  wmd::WarmInferArgs a = warm_args(c, c->last_lamda);
  const int64_t n_pairs = (int64_t)c->nq * c->N;
  const int64_t total = (int64_t)c->nq * c->nnz;
  constexpr int F = wmd::kWarmFeatures;
  if (o.standardize) {
    // z = (x - mean) / std per feature over every training sample.
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(1024, (total + 255) / 256));
    DevBuf<double> part;
    HIP_TRY(part.ensure((size_t)blocks * 2 * F));
    HIP_TRY(wmd::launch_warm_moments(a, part.p, blocks, c->stream));
    std::vector<double> hp((size_t)blocks * 2 * F);
    HIP_TRY(hipMemcpyAsync(hp.data(), part.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    part.release();
    float xm[F], xs[F];
    for (int k = 0; k < F; ++k) {
      double s1 = 0.0, s2 = 0.0;
      for (int b = 0; b < blocks; ++b) {
        s1 += hp[(size_t)b * 2 * F + k];
        s2 += hp[(size_t)b * 2 * F + F + k];
      }
      const double mean = total ? s1 / (double)total : 0.0;
      const double var = total ? std::max(0.0, s2 / (double)total - mean * mean) : 0.0;
      xm[k] = (float)mean;
      xs[k] = var > 1e-12 ? (float)(1.0 / std::sqrt(var)) : 1.f;
    }
    HIP_TRY(hipMemcpy(c->mlp_p.p + wmd::kWarmXm, xm, sizeof(xm), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->mlp_p.p + wmd::kWarmXs, xs, sizeof(xs), hipMemcpyHostToDevice));
  }
  // Sample layout per epoch: pairs in shuffled order, each pair's nonzeros
  // contiguous; sample = (query, nonzero), target = that nonzero's dual.
  std::vector<int64_t> order(n_pairs);
  for (int64_t k = 0; k < n_pairs; ++k) order[k] = k;
  std::vector<int32_t> s_q(total), s_pos(total);
  std::vector<int64_t> s_y(total);
  const int P = o.pairs_per_batch;
  const int64_t n_batches = (n_pairs + P - 1) / P;
  std::vector<int32_t> rel((size_t)n_batches * (P + 1));
  std::vector<int64_t> b_first(n_batches + 1);
  std::vector<int> b_groups(n_batches);
  DevBuf<int32_t> d_q, d_pos, d_rel;
  DevBuf<int64_t> d_y;
  DevBuf<float> gpart;
  DevBuf<double> lpart, eloss;
  HIP_TRY(gpart.ensure((size_t)P * wmd::kWarmParams));
  HIP_TRY(lpart.ensure((size_t)P));
  HIP_TRY(eloss.ensure((size_t)std::max(1, o.epochs)));
  HIP_TRY(hipMemsetAsync(eloss.p, 0, sizeof(double) * (size_t)std::max(1, o.epochs), c->stream));
  std::mt19937_64 rng(o.seed);
  int status = WMD_OK;
  for (int ep = 0; ep < o.epochs && status == WMD_OK; ++ep) {
    std::shuffle(order.begin(), order.end(), rng);
    int64_t at = 0;
    for (int64_t b = 0; b < n_batches; ++b) {
      b_first[b] = at;
      const int64_t p0 = b * P, p1 = std::min<int64_t>(n_pairs, p0 + P);
      b_groups[b] = (int)(p1 - p0);
      int32_t* rb = rel.data() + (size_t)b * (P + 1);
      rb[0] = 0;
      for (int64_t k = p0; k < p1; ++k) {
        const int64_t pr = order[k];
        const int Q = (int)(pr / c->N), j = (int)(pr % c->N);
        for (int32_t pos = c->h_doc_ptr[j]; pos < c->h_doc_ptr[j + 1]; ++pos) {
          s_q[at] = Q;
          s_pos[at] = pos;
          s_y[at] = (int64_t)Q * c->nnz + pos;
          ++at;
        }
        rb[k - p0 + 1] = (int32_t)(at - b_first[b]);
      }
      rc = check_train_batch(c, s_q.data() + b_first[b], s_pos.data() + b_first[b],
                             s_y.data() + b_first[b], at - b_first[b], rb, b_groups[b]);
      if (rc != WMD_OK) return rc;
    }
    b_first[n_batches] = at;
    HIP_TRY(d_q.upload(s_q.data(), (size_t)at, c->stream));
    HIP_TRY(d_pos.upload(s_pos.data(), (size_t)at, c->stream));
    HIP_TRY(d_y.upload(s_y.data(), (size_t)at, c->stream));
    HIP_TRY(d_rel.upload(rel.data(), rel.size(), c->stream));
    for (int64_t b = 0; b < n_batches; ++b) {
      const int64_t s0 = b_first[b], nb = b_first[b + 1] - s0;
      if (nb <= 0) continue;
      wmd::WarmTrainArgs t{};
      t.ctx = a;
      t.s_q = d_q.p + s0;
      t.s_pos = d_pos.p + s0;
      t.ysrc = c->duals.p;
      t.yidx = d_y.p + s0;
      t.gptr = d_rel.p + (size_t)b * (P + 1);
      t.groups = b_groups[b];
      t.n_total = nb;
      t.params = c->mlp_p.p;
      t.gpart = gpart.p;
      t.lpart = lpart.p;
      HIP_TRY(wmd::launch_warm_train(t, c->stream));
      ++c->adam_step;
      HIP_TRY(wmd::launch_warm_adam(c->mlp_p.p, gpart.p, t.groups, c->mlp_m.p, c->mlp_v.p,
                                    (float)o.lr, 0.9f, 0.999f, 1e-8f, c->adam_step,
                                    (float)o.weight_decay, lpart.p, eloss.p + ep, c->stream));
    }
    double el = 0.0;
    HIP_TRY(hipMemcpyAsync(&el, eloss.p + ep, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    el = at ? el / (double)at : 0.0;
    if (epoch_loss) epoch_loss[ep] = el;
    if (!std::isfinite(el)) {
      set_err("training diverged in epoch %d", ep);
      status = WMD_ERR_NONFINITE;
    }
  }
  d_q.release(); d_pos.release(); d_y.release(); d_rel.release();
  gpart.release(); lpart.release(); eloss.release();
  return status;
}

int wmd_mlp_loss_grad(int device, uint64_t n, const float* X, const float* y, uint64_t n_groups,
                      const int32_t* group_ptr, const float* params, double* loss, float* grad) {
  if (!X || !y || !group_ptr || !params || !loss || !grad || n == 0 || n_groups == 0 ||
      n > (1u << 24) || n_groups > (1u << 20)) {
    set_err("bad arguments");
    return WMD_ERR_ARG;
  }
  if (group_ptr[0] != 0 || (uint64_t)group_ptr[n_groups] != n) {
    set_err("group_ptr must run from 0 to n");
    return WMD_ERR_ARG;
  }
  for (uint64_t g = 0; g < n_groups; ++g)
    if (group_ptr[g + 1] < group_ptr[g]) {
      set_err("group_ptr decreases");
      return WMD_ERR_ARG;
    }
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) {
    set_err("device %d not present", device);
    return WMD_ERR_ARG;
  }
  HIP_TRY(hipSetDevice(device));
  constexpr int F = wmd::kWarmFeatures;
  DevBuf<float> dX, dy, dp, dgp, dg;
  DevBuf<int32_t> dgr;
  DevBuf<double> dlp, dl;
  hipStream_t st = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  hipError_t e = dX.upload(X, n * F, st);
  if (e == hipSuccess) e = dy.upload(y, n, st);
  if (e == hipSuccess) e = dp.upload(params, wmd::kWarmParams, st);
  if (e == hipSuccess) e = dgr.upload(group_ptr, n_groups + 1, st);
  if (e == hipSuccess) e = dgp.ensure((size_t)n_groups * wmd::kWarmParams);
  if (e == hipSuccess) e = dg.ensure(wmd::kWarmParams);
  if (e == hipSuccess) e = dlp.ensure(n_groups);
  if (e == hipSuccess) e = dl.ensure(1);
  if (e == hipSuccess) {
    wmd::WarmTrainArgs t{};
    t.X = dX.p;
    t.ysrc = dy.p;
    t.gptr = dgr.p;
    t.groups = (int)n_groups;
    t.n_total = (int64_t)n;
    t.params = dp.p;
    t.gpart = dgp.p;
    t.lpart = dlp.p;
    e = wmd::launch_warm_train(t, st);
  }
  if (e == hipSuccess) e = wmd::launch_warm_reduce(dgp.p, (int)n_groups, dlp.p, dg.p, dl.p, st);
  double lsum = 0.0;
  if (e == hipSuccess) e = hipMemcpyAsync(grad, dg.p, sizeof(float) * wmd::kWarmParams,
                                          hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&lsum, dl.p, sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  dX.release(); dy.release(); dp.release(); dgr.release(); dgp.release(); dg.release();
  dlp.release(); dl.release();
  (void)hipStreamDestroy(st);
  if (e != hipSuccess) {
    set_err("MLP loss/gradient failed: %s", hipGetErrorString(e));
    return WMD_ERR_HIP;
  }
  *loss = lsum / (double)n;
  return WMD_OK;
}

}  // extern "C"
