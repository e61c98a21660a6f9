/*
 * wmd.h -- C ABI of the MI355X Sinkhorn-Knopp Word Mover's Distance engine
 * (libwmd.so, HIP kernels for gfx950, one GPU).
 *
 * Reference interface replaced: IntelLabs/Optimized-Implementation-of-Word-
 * Movers-Distance, `void sinkhorn_wmd(const double* r, const double* vecs,
 * REAL* WMD, const REAL lamda, const int max_iter, const u64 data_vocab_size,
 * const u64 num_docs, const u64 word2vec_word_embedding_size)` at
 * sinkhorn_wmd_fused.cpp:112-114 (same signature at
 * sinkhorn_wmd_nonfused.cpp:114-116), which reads its V x N word-major CSR
 * from the global `c_csr` (sinkhorn_wmd_fused.cpp:13, common.h:38-46).  Here
 * the CSR is passed explicitly, every call is reentrant, the callee never
 * frees caller memory, and every entry point returns a status code.
 *
 * Conventions
 *   lamda   fused convention, K = exp(lamda * M) (sinkhorn_wmd_fused.cpp:217,
 *           common.h:26 ships lamda = -1): lamda must be < 0.  The nonfused
 *           entry keeps its file's convention K = exp(-M * lamda)
 *           (sinkhorn_wmd_nonfused.cpp:172-173), so there lamda must be > 0.
 *   CSR     V x N, row = word, u64 row_ptr[V+1], u64 col_inds[nnz] (doc ids),
 *           double vals[nnz] > 0 (the reference's reader rejects w <= 0,
 *           sinkhorn_wmd_fused.cpp:474-478).
 *   r       dense query histogram r[V] >= 0; its nonzeros are the query words
 *           (select_non_empty_entries, sinkhorn_wmd_fused.cpp:45-63).
 *   WMD     double[N] per query, row-major [query][doc] for batches.  A
 *           document with no words yields NaN exactly as the reference
 *           (x = 0, u = inf, inf * 0); the call then returns WMD_ERR_NONFINITE
 *           after writing every output.
 */
#ifndef WMD_H
#define WMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum wmd_status {
  WMD_OK = 0,
  WMD_ERR_ARG = -1,         /* null pointer, bad size, bad lamda, invalid CSR or query */
  WMD_ERR_HIP = -2,         /* HIP runtime error (no device, allocation, launch) */
  WMD_ERR_NONFINITE = -3,   /* outputs written, but some are NaN/inf */
  WMD_ERR_STATE = -4,       /* context call order: corpus or queries missing */
  WMD_ERR_UNSUPPORTED = -5  /* shape outside the selected variant's limits */
};

enum wmd_variant {
  /* one launch per tile shape; each (query, doc) tile stays in registers for
   * all iterations (default).  Documents longer than the tiles (160 words)
   * run in the same call in one more launch that iterates each of their
   * pairs inside the kernel, re-reading the kernel rows from memory (plain
   * steps from the uniform start; wmd_stats.streamed_docs counts them). */
  WMD_VARIANT_RESIDENT = 0,
  /* one fused SDDMM+SpMM launch per iteration (sinkhorn_wmd_fused.cpp:249-319) */
  WMD_VARIANT_FUSED = 1,
  /* SDDMM launch + SpMM launch per iteration (sinkhorn_wmd_nonfused.cpp:226-261) */
  WMD_VARIANT_NONFUSED = 2
};

enum wmd_domain {
  /* FP32 scalings u, v on the row-rescaled kernel exp(lamda (M - min_i M)) */
  WMD_DOMAIN_LINEAR = 0,
  /* FP32 base-2 log potentials with log-sum-exp steps: for large |lamda|,
   * where the kernel underflows FP32 (resident and fused variants) */
  WMD_DOMAIN_LOG = 1
};

enum wmd_cdist {
  /* FP32 direct differences sum_k (a_k - b_k)^2 on the vector ALUs */
  WMD_CDIST_FP32 = 0,
  /* |a|^2 + |b|^2 - 2 a.b with a.b on the matrix cores (three-way BF16 split,
   * FP32 accumulation), FP64 norms and cancellation, FP64 direct differences
   * for near-identical pairs */
  WMD_CDIST_MFMA = 1
};

enum wmd_warm {
  /* u = q for every pair (the reference's start, sinkhorn_wmd_fused.cpp:188-191) */
  WMD_WARM_UNIFORM = 0,
  /* v from the context's trained MLP; a pair whose residual does not decrease
   * over a 64-iteration window restarts once from the uniform start
   * (resident variant, linear domain) */
  WMD_WARM_LEARNED = 1,
  /* v from values given with wmd_context_set_warm (e.g. a previous solve's
   * duals); same stall fallback */
  WMD_WARM_GIVEN = 2
};

enum wmd_accel {
  /* plain Sinkhorn steps (or over-relaxed ones, see omega) */
  WMD_ACCEL_NONE = 0,
  /* Anderson acceleration with a history of one step on the centred base-2
   * potentials l = log2 u - mean(log2 u), in either domain: with g the plain
   * update of l (centred) and f = g - l, the next l is l + f - gamma (dl +
   * df), gamma = (df . f) / (df . df), dl and df the changes since the
   * previous step.  A
   * step whose marginal residual rises (or is not finite) is rejected and
   * replaced by the previous point's plain update.  A pair whose best
   * residual does not improve over a 64-iteration window continues with
   * plain steps.  Same fixed point and the same stop test as plain
   * Sinkhorn; needs tol > 0 (fixed-iteration runs reproduce the reference's
   * plain iterates); omega is ignored. */
  WMD_ACCEL_ANDERSON = 1
};

/* Reuse of distances across the queries of a batch (sinkhorn_wmd_fused.cpp:
 * 197-223 computes them per query).  In the shared form the distances of the
 * batch's distinct query words to the corpus words are computed once and
 * every query's row-rescaled kernel block is gathered from them; outputs are
 * bitwise equal to the per-query form. */
enum wmd_share {
  /* shared when the batch has at least 5 query words per 4 distinct ones,
   * no query is longer than 512 words and there are at most 16384 distinct
   * words */
  WMD_SHARE_AUTO = 0,
  WMD_SHARE_OFF = 1,   /* per-query distances */
  WMD_SHARE_ON = 2     /* shared whenever the limits above allow, else
                          WMD_ERR_UNSUPPORTED */
};

typedef struct wmd_options {
  int32_t max_iter;    /* iteration cap; the reference runs exactly 15 (:12) */
  int32_t variant;     /* enum wmd_variant */
  double tol;          /* 0: exactly max_iter iterations (reference behaviour);
                          > 0: each (query, doc) pair stops once its marginal
                          residual sum_i |u_i (K v)_i - r_i| < tol */
  int32_t check_every; /* streaming variants with tol > 0: iterations between
                          device-side scans that list the pairs still
                          iterating (later iterations walk that list); the
                          host reads the list's length every 8 scans (0 -> 8) */
  int32_t shape;       /* resident variant: -1 automatic, else a tile-shape id */
  int32_t summarize;   /* 1: fill wmd_stats iteration/residual summaries (one
                          device-to-host copy of per-pair counters); 0: stage
                          times only */
  int32_t domain;      /* enum wmd_domain */
  int32_t cdist;       /* enum wmd_cdist */
  int32_t warm_start;  /* enum wmd_warm */
  int32_t export_duals;/* 1: keep the converged log v of every (query, nonzero)
                          for wmd_context_duals / training (resident variant) */
  int32_t accel;       /* enum wmd_accel (resident variant, tol > 0; both domains) */
  double omega;        /* resident variant: over-relaxation of the u-update,
                          u <- u (r / (u K v))^omega, 1 = plain Sinkhorn, in
                          [1, 2).  The fixed point and the residual test are
                          unchanged (v is always the exact update); a pair whose
                          residual rises falls back to omega = 1. */
  int32_t graph;       /* resident variant: 1 (default) launches the solver
                          stage from a HIP graph captured on first use and
                          replayed while queries, tuning and options stay the
                          same; 0 launches the kernels directly */
  int32_t share;       /* enum wmd_share: distance reuse across a batch */
  int32_t reserved[2];
} wmd_options;

typedef struct wmd_stats {
  double ms_total;     /* device time of the whole run (HIP events) */
  double ms_cdist;     /* distance + exp stage */
  double ms_solve;     /* iterations + final pass */
  int64_t pairs;       /* queries x docs */
  int32_t iters_max;   /* largest iteration count over pairs */
  int32_t launches;    /* kernel launches (graph launches count their nodes) */
  double iters_mean;
  double resid_max;    /* largest final marginal residual over pairs */
  int32_t variant;     /* variant that ran */
  int32_t fallbacks;   /* warm-started pairs restarted from the uniform start */
  double ms_warm;      /* learned warm start: MLP inference stage */
  int32_t shared;      /* 1: the distance stage ran in the shared form */
  int32_t distinct;    /* distinct query words of the batch */
  int32_t streamed_docs;/* documents whose pairs ran outside the register
                          tiles: in resident runs those longer than 160 words
                          (plain steps from the uniform start, same stop
                          rule), in streaming runs every document */
  int32_t reserved[1];
} wmd_stats;

/* Default options: max_iter 15, resident variant, tol 0, automatic shape,
 * summarize 1, linear domain, matrix-core distance stage, uniform start,
 * omega 1. */
void wmd_options_default(wmd_options* opt);

/* Drop-in for the fused reference sinkhorn_wmd (sinkhorn_wmd_fused.cpp:112-114)
 * with its global CSR made explicit.  Runs on device 0 and blocks until WMD is
 * written.  lamda < 0. */
int wmd_sinkhorn(const double* r, const double* vecs, double* WMD, double lamda,
                 int max_iter, uint64_t data_vocab_size, uint64_t num_docs,
                 uint64_t word2vec_word_embedding_size, const uint64_t* row_ptr,
                 const uint64_t* col_inds, const double* vals);

/* Drop-in for the nonfused reference sinkhorn_wmd (sinkhorn_wmd_nonfused.cpp:
 * 114-116): same arguments, K = exp(-M * lamda) so lamda > 0; runs the
 * nonfused SDDMM/SpMM kernels. */
int wmd_sinkhorn_nonfused(const double* r, const double* vecs, double* WMD, double lamda,
                          int max_iter, uint64_t data_vocab_size, uint64_t num_docs,
                          uint64_t word2vec_word_embedding_size, const uint64_t* row_ptr,
                          const uint64_t* col_inds, const double* vals);

/* wmd_sinkhorn with explicit options and statistics (either may be NULL).
 * lamda in the fused convention. */
int wmd_sinkhorn_ex(const double* r, const double* vecs, double* WMD, double lamda,
                    uint64_t data_vocab_size, uint64_t num_docs,
                    uint64_t word2vec_word_embedding_size, const uint64_t* row_ptr,
                    const uint64_t* col_inds, const double* vals,
                    const wmd_options* opt, wmd_stats* stats);

/* ---- Resident-corpus context (no reference counterpart: the reference
 * re-reads and re-uploads nothing because it has no device).  A context owns
 * one HIP stream and the device copies of the corpus, queries and results;
 * it is not thread-safe, distinct contexts are independent. */
typedef struct wmd_context wmd_context;

int wmd_context_create(int device, wmd_context** out);
void wmd_context_destroy(wmd_context* ctx);

/* Upload embeddings (FP32 or FP64 host array V x D, exactly one non-NULL)
 * and the reference's V x N word-major CSR.  The corpus is converted once
 * into the device layout (doc-major CSR over words that occur). */
int wmd_context_set_corpus(wmd_context* ctx, const float* vecs_f32, const double* vecs_f64,
                           uint64_t V, uint64_t D, const uint64_t* row_ptr,
                           const uint64_t* col_inds, const double* vals, uint64_t num_docs);

/* Upload n_queries queries as CSR: q_ptr[n+1], q_words (vocabulary ids,
 * distinct per query), q_weights > 0. */
int wmd_context_set_queries(wmd_context* ctx, uint64_t n_queries, const int64_t* q_ptr,
                            const int64_t* q_words, const double* q_weights);

/* Solve every (query, doc) pair on the device; results stay on the device. */
int wmd_context_run(wmd_context* ctx, double lamda, const wmd_options* opt, wmd_stats* stats);

/* Allocation-free run of one query against the context's corpus (no
 * counterpart in the reference; for callers that cannot allocate device
 * memory: code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The variant of wmd_context_set_queries + wmd_context_run for a
 * single query whose q words (int32 vocabulary ids) and q weights (FP32, > 0)
 * are in device memory, writing wmd[N] (FP64) and optionally iters[N] and
 * resid[N] (may be NULL) to device memory.  Every per-query buffer (the
 * kernel block, row minima, matrix-core query planes, counters) lives in a
 * library-owned device scratch array of 128 MiB per GPU, held until the
 * process ends; a query whose buffers exceed it (Vt * 4 * ceil4(q) bytes for
 * the kernel block, Vt the corpus's distinct words) returns
 * WMD_ERR_UNSUPPORTED.  The call allocates nothing (no hipMalloc, hipFree or
 * allocating library routine) and blocks until the outputs are written.
 * Resident variant, uniform start, no dual export; the matrix-core distance
 * stage needs the corpus planes of an earlier wmd_context_run with it.  An
 * invalid word or weight is detected on the device and returns WMD_ERR_ARG.
 * All calls on one GPU share its scratch array: calls that overlap on one
 * GPU (from several threads or contexts) race and must be serialised by the
 * caller. */
int wmd_context_run_query_noalloc(wmd_context* ctx, int32_t q, const int32_t* q_words,
                                  const float* q_weights, double lamda, const wmd_options* opt,
                                  double* wmd, int32_t* iters, float* resid);

/* Copy results to the host: wmd[n_queries * N], iters and resid may be NULL. */
int wmd_context_result(wmd_context* ctx, double* wmd, int32_t* iters, float* resid);

/* Device pointer of the n_queries x N FP64 result (valid until the next
 * set_queries / destroy), for callers that keep results on the GPU. */
const double* wmd_context_device_result(const wmd_context* ctx);

/* The k nearest documents of every query of the last run (no reference
 * counterpart: the reference writes all N scores, sinkhorn_wmd_fused.cpp:
 * 614-622; retrieval is the use its README names): doc_ids[n_queries][k] and
 * wmd[n_queries][k], ascending distance, ties in document order, documents
 * without a distance (NaN) after every other.  1 <= k <= N. */
int wmd_context_topk(wmd_context* ctx, int32_t k, int32_t* doc_ids, double* wmd);

/* ---- Learned warm start (no reference counterpart; it replaces the
 * reference's uniform start x = 1/q, sinkhorn_wmd_fused.cpp:188-191).  A
 * 15 -> 32 -> 32 -> 1 MLP with ReLU hidden layers predicts, for
 * every (query, document word), the converged log v_e (row-rescaled kernel
 * convention) from norm features of the word's embedding, a document summary
 * and a query summary (feature k):
 *   0, 1    the word's embedding norm n and n^2
 *   2..6    document: c-weighted mean, std, min, max of its words' norms, log n_j
 *   7..11   query: r-weighted mean, std, min, max of its words' norms, log q
 *   12      the word's c-weighted mid-rank by norm within its document (cdf)
 *   13      T: the query's r-weighted norm quantile at cdf (numpy.interp over
 *           the query's sorted norms and their mid-ranks)
 *   14      |lamda| (sqrt(T^2 + n^2) - min_i M[word, i])
 * Parameters are one flat FP32 array: xm[15], xs[15] (inputs enter as
 * (x - xm) * xs), W1 [32][15] row-major, b1 [32], W2 [32][32], b2 [32],
 * w3 [32], b3 [1].  Inference runs the two hidden layers on the matrix cores
 * (FP16 operands, FP32 accumulation); training and wmd_mlp_loss_grad run in
 * FP32.  Targets are centred per (query, document) pair: the
 * scaling pair (a u, v / a) gives the same plan, so only differences of
 * log v within a document matter. */
#define WMD_MLP_HIDDEN 32
#define WMD_MLP_FEATURES 15

uint64_t wmd_mlp_param_count(void);
int wmd_context_mlp_set(wmd_context* ctx, const float* params, uint64_t n_params);
int wmd_context_mlp_get(wmd_context* ctx, float* params, uint64_t n_params);

typedef struct wmd_train_options {
  int32_t epochs;          /* passes over every (query, document) pair */
  int32_t pairs_per_batch; /* pairs per Adam step (at most 4096) */
  double lr;               /* Adam learning rate (beta 0.9 / 0.999, eps 1e-8) */
  double weight_decay;     /* L2 coefficient added to the gradient */
  uint64_t seed;           /* pair shuffling */
  int32_t standardize;     /* 1: set xm / xs to the training samples' per-feature
                              mean and 1 / std before the first step */
  int32_t reserved[7];
} wmd_train_options;

void wmd_train_options_default(wmd_train_options* opt);

/* Train the context's MLP on the GPU against the duals of the last run that
 * exported them (same corpus and queries): every (query, document) pair of
 * the current batch is one training group; one fused forward/backward kernel
 * and one Adam kernel per step.  epoch_loss[epochs] (may be NULL) receives
 * the mean centred squared error of each epoch. */
int wmd_context_mlp_train(wmd_context* ctx, const wmd_train_options* opt, double* epoch_loss);

/* The 15 features above for every (query, nonzero) of the current batch, as
 * the last run's row minima give them: out[n_queries][nnz][15], nonzeros in
 * the document-major order of wmd_context_duals. */
int wmd_context_warm_features(wmd_context* ctx, float* out);

/* The learned start's prediction for every (query, nonzero) of the current
 * batch with the context's network, from the last run's row minima, as the
 * inference kernel computes it (hidden layers on the matrix cores, FP16
 * operands, FP32 accumulation): out[n_queries][nnz] predicted log v, in the
 * layout of wmd_context_duals. */
int wmd_context_warm_predictions(wmd_context* ctx, float* out);

/* Caller-supplied starting log v (n_queries x nnz, the layout of
 * wmd_context_duals) for runs with warm_start = WMD_WARM_GIVEN. */
int wmd_context_set_warm(wmd_context* ctx, const float* log_v);

/* Converged log v of the last run with export_duals: n_queries x nnz, in the
 * document-major nonzero order (documents ascending, words ascending). */
int wmd_context_duals(wmd_context* ctx, float* out);

/* The training step's loss and gradient on the GPU for n samples of raw
 * features X [n][15] with targets y [n] in n_groups contiguous groups
 * (group_ptr [n_groups + 1]): loss = sum over groups of the squared
 * deviations of (prediction - y) from their group mean, over n; params and
 * grad use the flat layout above (xm / xs get zero gradient).  Stateless
 * (device `device`). */
int wmd_mlp_loss_grad(int device, uint64_t n, const float* X, const float* y, uint64_t n_groups,
                      const int32_t* group_ptr, const float* params, double* loss, float* grad);

/* Number of resident tile shapes and their limits (for tuners/tests). */
int wmd_num_shapes(void);
int wmd_shape_info(int shape, int32_t* waves, int32_t* cap_q, int32_t* cap_n);
/* Domains a shape runs in: bit 0 linear, bit 1 log (negative: bad shape). */
int wmd_shape_domains(int shape);

/* ---- Tile-shape autotuner (no reference counterpart).  The resident variant
 * runs each (query, document length class) group with one register tile
 * shape.  By default a query takes the smallest shape that holds its words;
 * the autotuner instead times, per document class and per group of queries
 * sharing that default, the default and up to five further shapes (capacity
 * at most twice the default's) on a sample of the context's pairs at `iters`
 * fixed iterations, and keeps the fastest for later runs with shape = -1 in
 * the options' domain.  It runs the distance stage first and leaves no valid
 * result (run again before wmd_context_result).  ms_out (may be NULL)
 * receives the tuning time.  Tables are [n_classes][max_q + 1] shape ids
 * (-1: smallest that fits) per domain, with n_classes and max_q from
 * wmd_tuning_dims; document class c holds the documents of 8 c + 1 to
 * 8 (c + 1) words (class 0 also the empty ones).  n_entries is the table's
 * length in int32 entries and must equal n_classes * (max_q + 1): a table
 * saved by a build with other dimensions is refused (WMD_ERR_ARG). */
int wmd_context_autotune(wmd_context* ctx, double lamda, const wmd_options* opt, int32_t iters,
                         double* ms_out);
int wmd_tuning_dims(int32_t* n_classes, int32_t* max_q);
int wmd_context_get_tuning(wmd_context* ctx, int32_t domain, int32_t* table,
                           uint64_t n_entries);
/* Load a table (e.g. saved from an earlier autotune); NULL restores the
 * default (n_entries is then ignored).  Entries must hold their q (and have
 * a log-domain form for the log domain). */
int wmd_context_set_tuning(wmd_context* ctx, int32_t domain, const int32_t* table,
                           uint64_t n_entries);

/* ---- The reference's file formats (host only, no device), as read and
 * written by its programs' main (sinkhorn_wmd_fused.cpp:415-624):
 *   mat.mtx   "rows cols nnz", then nnz lines "word doc weight", 1-based ids,
 *             weight > 0 (:427-480) -> the V x N word-major CSR, each row in
 *             file order like the reference's counting sort (:494-506)
 *   vecs.out  one line per word of D space-separated values (:510-545)
 *   r.out     V values >= 0, one per line (:553-586)
 *   scores    one line per document, 9 significant digits (:614-622)
 * Invalid content returns WMD_ERR_ARG (the reference prints and exits). */
int wmd_mtx_header(const char* path, uint64_t* V, uint64_t* N, uint64_t* nnz);
int wmd_read_mtx(const char* path, uint64_t V, uint64_t N, uint64_t nnz, uint64_t* row_ptr,
                 uint64_t* col_inds, double* vals);
int wmd_vecs_dims(const char* path, uint64_t* V, uint64_t* D);
int wmd_read_vecs(const char* path, uint64_t V, uint64_t D, double* vecs);
int wmd_read_r(const char* path, uint64_t V, double* r);
int wmd_write_scores(const char* path, const double* wmd, uint64_t N);
/* Writers of the three input formats (no reference counterpart: they make
 * inputs for the reference's programs and for wmd_b), values with 17
 * significant digits so they read back exactly. */
int wmd_write_mtx(const char* path, uint64_t V, uint64_t N, const uint64_t* row_ptr,
                  const uint64_t* col_inds, const double* vals);
int wmd_write_vecs(const char* path, uint64_t V, uint64_t D, const double* vecs);
int wmd_write_r(const char* path, uint64_t V, const double* r);

/* ---- Corpus preparation from raw text (prep.py:17-50 load_vecs + docs2mat,
 * prep.py:88-91 reading the document file).  Host code; no device needed. */
enum wmd_text_format {
  /* comma-separated records ('"' quoting, '""' escapes); a document is the
   * second field, one space, the third field (DBpedia "class,title,content",
   * prep.py:89-90) */
  WMD_TEXT_CSV = 0,
  /* one document per line ('\n' or "\r\n" endings) */
  WMD_TEXT_LINES = 1
};
enum wmd_text_which { WMD_TEXT_TOKENS = 0, WMD_TEXT_DOCS = 1 };
typedef struct wmd_prepared wmd_prepared;
/* Reads the first n_toks embeddings of a fastText .vec file (header "count
 * dim"; a token repeated later maps to its last line) and the documents of
 * text_path; keeps the documents with at least min_toks distinct vocabulary
 * tokens (tokens split on single spaces); builds the V x N word-major CSR
 * with weight 1/len(document) per distinct token, documents ascending in each
 * row (the matrix prep.py:50 saves).  *out is released by wmd_prepared_free. */
int wmd_prepare_corpus(const char* vec_path, const char* text_path, int text_format,
                       uint64_t n_toks, uint64_t min_toks, wmd_prepared** out);
int wmd_prepared_dims(const wmd_prepared* p, uint64_t* V, uint64_t* D, uint64_t* N,
                      uint64_t* nnz);
/* Copies out V x D embeddings (row-major), the CSR (row_ptr V+1, col_inds and
 * vals nnz), per kept document its record index in text_path (records
 * counted from 0; blank lines are not records in the CSV format), and N+1
 * byte offsets into the WMD_TEXT_DOCS text: document j is the bytes
 * [doc_offsets[j], doc_offsets[j+1] - 1) (a document may itself hold line
 * breaks).  Any output may be null. */
int wmd_prepared_copy(const wmd_prepared* p, double* vecs, uint64_t* row_ptr,
                      uint64_t* col_inds, double* vals, uint64_t* kept_records,
                      uint64_t* doc_offsets);
/* The tokens (which = WMD_TEXT_TOKENS) or the kept documents' text
 * (WMD_TEXT_DOCS), '\n'-joined (prep.py:98-99); *needed = its byte length;
 * buf may be null to ask for the length only. */
int wmd_prepared_text(const wmd_prepared* p, int which, char* buf, uint64_t cap,
                      uint64_t* needed);
void wmd_prepared_free(wmd_prepared* p);

/* Human-readable description of the last error on this thread. */
const char* wmd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* WMD_H */
