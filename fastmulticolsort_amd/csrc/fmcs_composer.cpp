// fmcs_composer.cpp -- the reference's composer API on the host, driving
// the GPU through the C ABI.
//
// Column / SpecialiedColumn<T> (src/column.h:12-63) hold device banks;
// Composer (src/composer.h:20-176) runs the round loop; ChainComposer
// (src/chain_composer.h) is the concrete composer; StitchComposer
// (src/stitch_composer.h:13-252) massages columns into new banks and runs
// one plan.  Times are HIP-event milliseconds on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/fmcs.h"
#include "../../include/fmcs_plan.h"
#include "fmcs_internal.h"

namespace fmcs {
namespace host {

using internal::scratch;
using internal::set_error;
using internal::stream_of;

// construct_entry_t (src/types.h:42-46)
struct construct_entry_t {
  uint32_t finalIdx;
  uint32_t start;
  uint32_t end;
};

// setting_t's timing fields (src/types.h:198-236), in milliseconds.
struct setting_t {
  double time_tupleReconstruct = 0;
  double time_orderGroupInfo = 0;
  double time_multipass_sort = 0;
  double time_firstpass_sort = 0;
  double time_extractGrop = 0;
  double time_stitching = 0;
};

// compose_params_t (src/types.h:169-186), the fields this path uses.
struct compose_params_t {
  std::vector<uint32_t> bitwidth;
  std::vector<uint8_t> descending;  // per column, empty: all ascending (the reference: ascending only)
  uint32_t ordered = 1;
  uint32_t hash_type = 1;
  int mode = FMCS_MODE_BATCHED;
};

class Column {
 public:
  virtual ~Column() {}
  virtual void* GetColumn() const = 0;
  virtual uint32_t GetWidth() const = 0;
};

template <class T>
class SpecialiedColumn : public Column {
 public:
  explicit SpecialiedColumn(T* values) : values_(values) {}
  void* GetColumn() const override { return values_; }
  uint32_t GetWidth() const override { return sizeof(T); }

 private:
  T* values_;
};

std::unique_ptr<Column> make_column(uint32_t bank, void* p) {
  switch (bank) {
    case 1: return std::unique_ptr<Column>(new SpecialiedColumn<uint8_t>((uint8_t*)p));
    case 2: return std::unique_ptr<Column>(new SpecialiedColumn<uint16_t>((uint16_t*)p));
    case 4: return std::unique_ptr<Column>(new SpecialiedColumn<uint32_t>((uint32_t*)p));
    default: return std::unique_ptr<Column>(new SpecialiedColumn<uint64_t>((uint64_t*)p));
  }
}

class Composer {
 public:
  Composer(fmcs_ctx* ctx, uint32_t ncolumns, uint64_t nrows, compose_params_t params, std::vector<Column*> cols)
      : ctx_(ctx), num_columns_(ncolumns), num_rows_(nrows), params_(std::move(params)), cols_(std::move(cols)) {}
  virtual ~Composer() {}

  // Composer::SortAllColumns -> multiRoundSorting (src/composer.cpp:18-27, :48-135)
  int SortAllColumns(uint32_t* order_d, uint32_t* groups_d, uint32_t* ngroups) {
    std::vector<const void*> p(num_columns_);
    std::vector<uint32_t> banks(num_columns_);
    for (uint32_t i = 0; i < num_columns_; ++i) {
      p[i] = cols_[i]->GetColumn();
      banks[i] = cols_[i]->GetWidth();
    }
    int rc = fmcs_multi_round_sort(ctx_, num_columns_, p.data(), banks.data(), params_.bitwidth.data(), num_rows_,
                                   order_d, groups_d, params_.mode, ngroups);
    if (rc != FMCS_OK) return rc;
    double b[4];
    fmcs_last_breakdown(ctx_, b);
    setting_.time_tupleReconstruct = b[0];
    setting_.time_firstpass_sort = b[1];
    setting_.time_multipass_sort = b[1] + b[2];
    setting_.time_extractGrop = b[3];
    setting_.time_orderGroupInfo = b[3];
    return FMCS_OK;
  }

  // ChainComposer::HashAllColumns (src/chain_composer.cpp:16-25)
  int HashAllColumns(uint32_t* order_d, uint32_t* gid_d, uint32_t* ngroups) {
    std::vector<const void*> p(num_columns_);
    std::vector<uint32_t> banks(num_columns_);
    for (uint32_t i = 0; i < num_columns_; ++i) {
      p[i] = cols_[i]->GetColumn();
      banks[i] = cols_[i]->GetWidth();
    }
    int rc = fmcs_multi_round_hash(ctx_, num_columns_, p.data(), banks.data(), num_rows_, (int)params_.hash_type,
                                   order_d, gid_d, ngroups);
    if (rc != FMCS_OK) return rc;
    double b[4];
    fmcs_last_breakdown(ctx_, b);
    setting_.time_multipass_sort = b[1];
    return FMCS_OK;
  }

  setting_t getSetting() const { return setting_; }

 protected:
  fmcs_ctx* ctx_;
  uint32_t num_columns_;
  uint64_t num_rows_;
  compose_params_t params_;
  std::vector<Column*> cols_;
  setting_t setting_;
};

class ChainComposer : public Composer {
 public:
  using Composer::Composer;
};

class StitchComposer : public Composer {
 public:
  using Composer::Composer;

  // Step a of run_single_plan (src/stitch_composer.cpp:571-612).
  int construct(const uint32_t* columnOrder, const std::vector<uint32_t>& split,
                std::vector<fmcs_entry>& entries) {
    entries.resize(64 * 64);
    uint32_t ne = 0;
    int rc = fmcs_construct_entries(num_columns_, params_.bitwidth.data(), columnOrder, (uint32_t)split.size(),
                                    split.data(), entries.data(), (uint32_t)entries.size(), &ne);
    if (rc != FMCS_OK) return set_error(ctx_, rc, "invalid split for these columns");
    entries.resize(ne);
    return FMCS_OK;
  }

  // StitchComposer::run_single_plan / run_single_plan_hash
  // (src/stitch_composer.cpp:555-806, :808-1079): b1 allocate the new banks,
  // b2 stitch them (timed as breakdown[0]), c run a ChainComposer on them.
  int run_single_plan(const uint32_t* columnOrder, const std::vector<uint32_t>& split, bool hash,
                      uint32_t* order_d, uint32_t* groups_d, double* breakdown, double* total_ms,
                      uint32_t* ngroups) {
    const uint32_t nn = (uint32_t)split.size();
    if (nn == 0 || nn > 16) return set_error(ctx_, FMCS_E_ARG, "1..16 rounds supported");
    std::vector<fmcs_entry> entries;
    int rc = construct(columnOrder, split, entries);
    if (rc != FMCS_OK) return rc;
    std::vector<void*> dst(nn);
    std::vector<uint32_t> dbank(nn);
    for (uint32_t c = 0; c < nn; ++c) {
      dbank[c] = fmcs_bank_bytes(split[c], hash ? 1 : 0);
      rc = scratch(ctx_, (int)c, (num_rows_ + 16) * dbank[c], &dst[c]);
      if (rc != FMCS_OK) return rc;
    }
    std::vector<const void*> src(num_columns_);
    std::vector<uint32_t> sbank(num_columns_);
    for (uint32_t i = 0; i < num_columns_; ++i) {
      src[i] = cols_[i]->GetColumn();
      sbank[i] = cols_[i]->GetWidth();
    }
    hipStream_t st = stream_of(ctx_);
    hipEvent_t e0, e1, e2;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess || hipEventCreate(&e2) != hipSuccess)
      return set_error(ctx_, FMCS_E_HIP, "event creation failed");
    (void)hipEventRecord(e0, st);
    rc = fmcs_stitch_dir(ctx_, num_rows_, num_columns_, src.data(), sbank.data(), params_.bitwidth.data(),
                         (uint32_t)entries.size(), entries.data(), nn, dst.data(), dbank.data(),
                         params_.descending.empty() ? nullptr : params_.descending.data());
    (void)hipEventRecord(e1, st);
    std::vector<std::unique_ptr<Column>> owned;
    std::vector<Column*> newcols;
    for (uint32_t c = 0; c < nn; ++c) {
      owned.push_back(make_column(dbank[c], dst[c]));
      newcols.push_back(owned.back().get());
    }
    compose_params_t p2;
    p2.bitwidth = split;
    p2.ordered = params_.ordered;
    p2.hash_type = params_.hash_type;
    p2.mode = params_.mode;
    ChainComposer chain(ctx_, nn, num_rows_, p2, newcols);
    if (rc == FMCS_OK)
      rc = hash ? chain.HashAllColumns(order_d, groups_d, ngroups) : chain.SortAllColumns(order_d, groups_d, ngroups);
    (void)hipEventRecord(e2, st);
    (void)hipEventSynchronize(e2);
    float ms_stitch = 0, ms_all = 0;
    (void)hipEventElapsedTime(&ms_stitch, e0, e1);
    (void)hipEventElapsedTime(&ms_all, e0, e2);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (rc != FMCS_OK) return rc;
    setting_t u = chain.getSetting();
    if (breakdown) {
      breakdown[0] = ms_stitch;
      breakdown[1] = u.time_tupleReconstruct;
      breakdown[2] = u.time_orderGroupInfo;
      breakdown[3] = u.time_multipass_sort;
      breakdown[4] = u.time_firstpass_sort;
      breakdown[5] = u.time_extractGrop;
    }
    if (total_ms) *total_ms = ms_all;
    return FMCS_OK;
  }
};

// ---- calibration -----------------------------------------------------------

double time_ms(fmcs_ctx* ctx, int reps, int (*fn)(fmcs_ctx*, void*), void* arg) {
  hipStream_t st = stream_of(ctx);
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  fn(ctx, arg);  // warm-up
  (void)hipEventRecord(a, st);
  for (int i = 0; i < reps; ++i)
    if (fn(ctx, arg) != FMCS_OK) break;
  (void)hipEventRecord(b, st);
  (void)hipEventSynchronize(b);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

// The least of `trials` runs of time_ms (the per-row constants: large
// sorts, where jitter only ever adds time).  The fixed costs of small sorts
// are means over many runs instead: plans pay the typical launch cost, and a
// minimum would price it too low.
double time_ms_best(fmcs_ctx* ctx, int trials, int reps, int (*fn)(fmcs_ctx*, void*), void* arg) {
  double best = 1e300;
  for (int t = 0; t < trials; ++t) best = std::min(best, time_ms(ctx, reps, fn, arg));
  return best;
}

struct CalArgs {
  uint32_t bank, bits;
  uint64_t n;
  void *keys, *keys0, *vals2;
  uint32_t *oids, *oids0, *perm, *off, *grp;
  uint32_t nseg;
};

int reset_pairs(fmcs_ctx* ctx, CalArgs* a) {
  hipStream_t st = stream_of(ctx);
  if (hipMemcpyAsync(a->keys, a->keys0, a->n * a->bank, hipMemcpyDeviceToDevice, st) != hipSuccess) return FMCS_E_HIP;
  if (hipMemcpyAsync(a->oids, a->oids0, a->n * 4, hipMemcpyDeviceToDevice, st) != hipSuccess) return FMCS_E_HIP;
  return FMCS_OK;
}

}  // namespace host
}  // namespace fmcs

using namespace fmcs::host;

namespace {

// Checks (and, if the context stages host buffers, stages) the column banks
// and the order / group outputs of a plan run.
int stage_plan_args(fmcs::internal::Stager& st, uint32_t ncols, const void* const* cols_d, const uint32_t* banks,
                    uint64_t nrows, std::vector<const void*>& cols, uint32_t*& order_d, uint32_t*& groups_d) {
  using fmcs::internal::Stager;
  cols.assign(cols_d, cols_d + ncols);
  for (uint32_t i = 0; i < ncols; ++i) {
    if (banks[i] != 2 && banks[i] != 4 && banks[i] != 8) return FMCS_E_ARG;
    int rc = st.arg(cols[i], nrows * banks[i], Stager::IN, "cols_d");
    if (rc != FMCS_OK) return rc;
  }
  int rc = st.arg(order_d, nrows * 4, Stager::OUT, "order_d");
  if (rc == FMCS_OK) rc = st.arg(groups_d, nrows * 4, Stager::OUT, "groups_d");
  return rc;
}

}  // namespace

extern "C" {

int fmcs_run_plan(fmcs_ctx* ctx, uint32_t ncols, const void* const* cols_d, const uint32_t* banks, const uint32_t* bw,
                  const uint32_t* col_order, uint32_t nsplit, const uint32_t* split, int mode, uint64_t nrows,
                  uint32_t* order_d, uint32_t* groups_d, double* ms6, double* total_ms, uint32_t* ngroups) {
  return fmcs_run_plan_dir(ctx, ncols, cols_d, banks, bw, nullptr, col_order, nsplit, split, mode, nrows, order_d,
                           groups_d, ms6, total_ms, ngroups);
}

int fmcs_run_plan_dir(fmcs_ctx* ctx, uint32_t ncols, const void* const* cols_d, const uint32_t* banks,
                      const uint32_t* bw, const uint8_t* descending, const uint32_t* col_order, uint32_t nsplit,
                      const uint32_t* split, int mode, uint64_t nrows, uint32_t* order_d, uint32_t* groups_d,
                      double* ms6, double* total_ms, uint32_t* ngroups) {
  if (!ctx || !cols_d || !banks || !bw || !col_order || !split || !order_d || ncols == 0 || ncols > 64)
    return FMCS_E_ARG;
  fmcs::internal::Stager st(ctx);
  std::vector<const void*> cv;
  int rc = stage_plan_args(st, ncols, cols_d, banks, nrows, cv, order_d, groups_d);
  if (rc != FMCS_OK) return rc;
  std::vector<std::unique_ptr<Column>> owned;
  std::vector<Column*> cols;
  for (uint32_t i = 0; i < ncols; ++i) {
    owned.push_back(make_column(banks[i], const_cast<void*>(cv[i])));
    cols.push_back(owned.back().get());
  }
  compose_params_t p;
  p.bitwidth.assign(bw, bw + ncols);
  if (descending) p.descending.assign(descending, descending + ncols);
  p.mode = mode;
  StitchComposer sc(ctx, ncols, nrows, p, cols);
  std::vector<uint32_t> sp(split, split + nsplit);
  rc = sc.run_single_plan(col_order, sp, false, order_d, groups_d, ms6, total_ms, ngroups);
  return rc != FMCS_OK ? rc : st.finish();
}

int fmcs_optimize_and_run(fmcs_ctx* ctx, const fmcs_cost_model* m, uint32_t ncols, const void* const* cols_d,
                          const uint32_t* banks, const uint32_t* bw, const uint64_t* card, int ordered, int mode,
                          uint64_t nrows, uint32_t* order_d, uint32_t* groups_d, uint32_t* best_order,
                          uint32_t* best_split, uint32_t* best_nsplit, double* opt_ms, double* ms6, double* total_ms,
                          uint32_t* ngroups) {
  if (!ctx || !cols_d || !banks || !bw || !order_d || !best_order || !best_split || !best_nsplit || ncols == 0 ||
      ncols > 64)
    return FMCS_E_ARG;
  fmcs::internal::Stager st(ctx);
  std::vector<const void*> staged;
  int src = stage_plan_args(st, ncols, cols_d, banks, nrows, staged, order_d, groups_d);
  if (src != FMCS_OK) return src;
  cols_d = staged.data();
  auto t0 = std::chrono::steady_clock::now();
  fmcs_cost_model dm;
  if (!m) {
    fmcs_cost_model_default(&dm);
    m = &dm;
  }
  // base column cardinalities (StitchComposer's constructor,
  // src/stitch_composer.h:22-37): given, or counted on the device
  std::vector<uint64_t> cv(ncols);
  for (uint32_t c = 0; c < ncols; ++c) {
    if (card) {
      cv[c] = card[c];
    } else {
      int rc = fmcs_distinct_count(ctx, banks[c], nrows, cols_d[c], &cv[c]);
      if (rc != FMCS_OK) return rc;
    }
    cv[c] = std::max<uint64_t>(cv[c], 1);
  }
  // the search may return up to ceil((2S - 2) / 16) rounds; the executor
  // runs at most 16
  uint32_t S = 0;
  for (uint32_t c = 0; c < ncols; ++c) S += bw[c];
  std::vector<uint32_t> ord(ncols), split((2 * S) / 16 + 64);
  uint32_t ns = 0;
  double cost = 0;
  int rc = fmcs_optimize(m, nrows, ncols, bw, cv.data(), ordered, ord.data(), split.data(), &ns, &cost, nullptr);
  if (rc != FMCS_OK) return set_error(ctx, rc, "plan search failed");
  if (ns > 16) return set_error(ctx, FMCS_E_UNSUPPORTED, "best plan has more than 16 rounds");
  if (opt_ms) *opt_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::copy(ord.begin(), ord.end(), best_order);
  std::copy(split.begin(), split.begin() + ns, best_split);
  *best_nsplit = ns;
  rc = fmcs_run_plan(ctx, ncols, cols_d, banks, bw, ord.data(), ns, split.data(), mode, nrows, order_d, groups_d, ms6,
                     total_ms, ngroups);
  return rc != FMCS_OK ? rc : st.finish();
}

int fmcs_run_plan_hash(fmcs_ctx* ctx, uint32_t ncols, const void* const* cols_d, const uint32_t* banks,
                       const uint32_t* bw, const uint32_t* col_order, uint32_t nsplit, const uint32_t* split,
                       int hash_type, uint64_t nrows, uint32_t* order_d, uint32_t* gid_d, double* total_ms,
                       uint32_t* ngroups) {
  if (!ctx || !cols_d || !banks || !bw || !col_order || !split || !order_d || !gid_d || ncols == 0 || ncols > 64)
    return FMCS_E_ARG;
  fmcs::internal::Stager st(ctx);
  std::vector<const void*> cv;
  int rc = stage_plan_args(st, ncols, cols_d, banks, nrows, cv, order_d, gid_d);
  if (rc != FMCS_OK) return rc;
  std::vector<std::unique_ptr<Column>> owned;
  std::vector<Column*> cols;
  for (uint32_t i = 0; i < ncols; ++i) {
    owned.push_back(make_column(banks[i], const_cast<void*>(cv[i])));
    cols.push_back(owned.back().get());
  }
  compose_params_t p;
  p.bitwidth.assign(bw, bw + ncols);
  p.hash_type = (uint32_t)hash_type;
  StitchComposer sc(ctx, ncols, nrows, p, cols);
  std::vector<uint32_t> sp(split, split + nsplit);
  rc = sc.run_single_plan(col_order, sp, true, order_d, gid_d, nullptr, total_ms, ngroups);
  return rc != FMCS_OK ? rc : st.finish();
}

int fmcs_calibrate(fmcs_ctx* ctx, uint64_t nrows, fmcs_cost_model* out) {
  if (!ctx || !out || nrows < 4096 || nrows >= (1ull << 30)) return FMCS_E_ARG;
  fmcs_cost_model_default(out);
  fmcs_tuning tune;
  if (fmcs_ctx_get_tuning(ctx, &tune) == FMCS_OK) out->carry_next = tune.carry_next;
  const uint64_t n = nrows;
  CalArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = n;
  int rc;
  void *k, *k0, *v2;
  uint32_t *o, *o0, *perm, *off;
  if ((rc = scratch(ctx, 16, n * 8 + 64, &k)) || (rc = scratch(ctx, 17, n * 8 + 64, &k0)) ||
      (rc = scratch(ctx, 18, n * 8 + 64, &v2)) || (rc = scratch(ctx, 19, n * 4 + 64, (void**)&o)) ||
      (rc = scratch(ctx, 20, n * 4 + 64, (void**)&o0)) || (rc = scratch(ctx, 21, n * 4 + 64, (void**)&perm)) ||
      (rc = scratch(ctx, 22, n * 4 + 64, (void**)&off)))
    return rc;
  a.keys = k; a.keys0 = k0; a.vals2 = v2; a.oids = o; a.oids0 = o0; a.perm = perm; a.off = off;
  // a random permutation: sort random keys with an iota payload
  if ((rc = fmcs::internal::fill_random(ctx, k0, 8, n, 12345, 64))) return rc;
  if ((rc = fmcs::internal::uniform_offsets(ctx, o0, (uint32_t)n - 1, n, 1))) return rc;  // o0 = 0..n-1
  if (hipMemcpyAsync(perm, o0, n * 4, hipMemcpyDeviceToDevice, stream_of(ctx)) != hipSuccess) return FMCS_E_HIP;
  if ((rc = fmcs_sort_pairs(ctx, 8, 64, n, k0, perm))) return rc;
  // Fixed costs (launches, list set-up) of the device-driven plans, from
  // sorts of a few rows (means of 60 runs each): per LSD pass (a 40-bit against a 16-bit first
  // round: 5 passes against 2), per level of a later round (a 48-bit against
  // a 16-bit second round: 7 levels against 3) and per round (two rounds
  // against one, less the second round's levels).
  {
    a.n = 4096;
    if ((rc = fmcs::internal::fill_random(ctx, k0, 8, a.n, 31, 40)) || (rc = fmcs::internal::fill_random(ctx, k, 8, a.n, 32, 48)) ||
        (rc = fmcs::internal::fill_random(ctx, v2, 8, a.n, 33, 16)))
      return rc;
    // u64 banks: keys0 40-bit, keys 48-bit, vals2 16-bit values
    struct Run {
      static int go(fmcs_ctx* c, CalArgs* q, uint32_t rounds, const void* c0, uint32_t w0, const void* c1, uint32_t w1) {
        const void* cols[2] = {c0, c1};
        const uint32_t bk[2] = {8, 8}, bw[2] = {w0, w1};
        return fmcs_multi_round_sort(c, rounds, cols, bk, bw, q->n, q->oids, nullptr, FMCS_MODE_DEVICE, nullptr);
      }
    };
    double t20 = time_ms(ctx, 60, [](fmcs_ctx* c, void* x) {
      CalArgs* q = (CalArgs*)x;
      return Run::go(c, q, 1, q->vals2, 16, nullptr, 0);
    }, &a);
    double t40 = time_ms(ctx, 60, [](fmcs_ctx* c, void* x) {
      CalArgs* q = (CalArgs*)x;
      return Run::go(c, q, 1, q->keys0, 40, nullptr, 0);
    }, &a);
    double t2a = time_ms(ctx, 60, [](fmcs_ctx* c, void* x) {
      CalArgs* q = (CalArgs*)x;
      return Run::go(c, q, 2, q->vals2, 16, q->vals2, 16);
    }, &a);
    double t2b = time_ms(ctx, 60, [](fmcs_ctx* c, void* x) {
      CalArgs* q = (CalArgs*)x;
      return Run::go(c, q, 2, q->vals2, 16, q->keys, 48);
    }, &a);
    out->pass_overhead_ms = std::max(0.001, (t40 - t20) / 3.0);  // 5 passes against 2
    out->level_overhead_ms = std::max(0.002, (t2b - t2a) / 4.0);  // 7 levels against 3
    out->round_overhead_ms = std::max(0.005, t2a - t20 - 3.0 * out->level_overhead_ms);
    if ((rc = fmcs_ctx_sync(ctx))) return rc;
    a.n = n;
  }
  const uint32_t banks[3] = {2, 4, 8}, bits[3] = {16, 32, 64};
  for (int b = 0; b < 3; ++b) {
    a.bank = banks[b];
    a.bits = bits[b];
    // one LSD pass as the plans run it: device-driven single-round sorts of
    // three 8-bit digits against one (two against one in a u16 bank), so the
    // fixed costs of a call cancel (an eager sort would add its host
    // synchronisation to every pass)
    auto one_round = [](fmcs_ctx* c, void* x) {
      CalArgs* q = (CalArgs*)x;
      const void* cols[1] = {q->keys0};
      const uint32_t bk[1] = {q->bank}, w[1] = {q->bits};
      return fmcs_multi_round_sort(c, 1, cols, bk, w, q->n, q->oids, nullptr, FMCS_MODE_DEVICE, nullptr);
    };
    const uint32_t wide = banks[b] == 2 ? 16u : 24u;
    a.bits = 8;
    if ((rc = fmcs::internal::fill_random(ctx, k0, a.bank, n, 777 + b, 8))) return rc;
    const double t_one = time_ms_best(ctx, 3, 3, one_round, &a);
    a.bits = wide;
    if ((rc = fmcs::internal::fill_random(ctx, k0, a.bank, n, 778 + b, wide))) return rc;
    const double t_wide = time_ms_best(ctx, 3, 3, one_round, &a);
    a.bits = bits[b];
    out->radix_pass_ns[b] = std::max(1e-6, (t_wide - t_one) / (wide / 8 - 1)) * 1e6 / n;
    double t_copy = time_ms_best(ctx, 3, 3, [](fmcs_ctx* c, void* x) { return reset_pairs(c, (CalArgs*)x); }, &a);
    if ((rc = fmcs::internal::fill_random(ctx, k0, a.bank, n, 999 + b, a.bits))) return rc;
    double t_g = time_ms_best(ctx, 3, 3, [](fmcs_ctx* c, void* x) {
      CalArgs* q = (CalArgs*)x;
      return fmcs_gather(c, q->bank, q->n, q->keys0, q->perm, q->keys);
    }, &a);
    out->gather_ns[b] = t_g * 1e6 / n;
    // boundary over sorted keys
    if ((rc = reset_pairs(ctx, &a)) || (rc = fmcs_sort_pairs(ctx, a.bank, a.bits, n, a.keys, a.oids))) return rc;
    double t_b = time_ms_best(ctx, 3, 3, [](fmcs_ctx* c, void* x) {
      CalArgs* q = (CalArgs*)x;
      uint32_t ng;
      return fmcs_extract_group_info(c, q->bank, q->n, q->keys, q->oids, (uint32_t*)q->vals2, nullptr, &ng);
    }, &a);
    out->boundary_ns[b] = t_b * 1e6 / n;
    // the fixed cost of a device-driven segmented sort of these keys (the
    // launches of its level schedule), from a 4096-row sort in one-row
    // segments; the plan estimate adds it per level itself
    double t_fix;
    {
      a.n = 4096;
      a.nseg = 4096;
      if ((rc = fmcs::internal::uniform_offsets(ctx, off, a.nseg, a.n, 1u))) return rc;
      t_fix = time_ms_best(ctx, 3, 3, [](fmcs_ctx* c, void* x) {
        CalArgs* q = (CalArgs*)x;
        int r = reset_pairs(c, q);
        return r ? r : fmcs_segmented_sort(c, q->bank, q->bits, q->n, q->keys, q->oids, q->off, q->nseg,
                                           FMCS_MODE_DEVICE, nullptr);
      }, &a);
      a.n = n;
    }
    // segmented sorts over uniform groups of 2^i rows
    for (int i = 0; i < FMCS_COST_GBUCKETS; ++i) {
      uint64_t g = 1ull << i;
      if (g > n) {
        out->segsort_ns[b][i] = out->segsort_ns[b][i - 1];
        continue;
      }
      a.nseg = (uint32_t)((n + g - 1) / g);
      if ((rc = fmcs::internal::uniform_offsets(ctx, off, a.nseg, n, (uint32_t)g))) return rc;
      double t = time_ms_best(ctx, 3, 2, [](fmcs_ctx* c, void* x) {
        CalArgs* q = (CalArgs*)x;
        int r = reset_pairs(c, q);
        return r ? r : fmcs_segmented_sort(c, q->bank, q->bits, q->n, q->keys, q->oids, q->off, q->nseg,
                                           FMCS_MODE_DEVICE, nullptr);
      }, &a);
      out->segsort_ns[b][i] = std::max(1e-6, t - t_copy - t_fix) * 1e6 / n;
    }
  }
  // stitch: two u32 sources into one u64 bank
  {
    a.bank = 4;
    if ((rc = fmcs::internal::fill_random(ctx, k0, 4, n, 5, 20)) || (rc = fmcs::internal::fill_random(ctx, k, 4, n, 6, 20)))
      return rc;
    double t_s = time_ms_best(ctx, 3, 3, [](fmcs_ctx* c, void* x) {
      CalArgs* q = (CalArgs*)x;
      const void* src[2] = {q->keys0, q->keys};
      uint32_t sb[2] = {4, 4}, sw[2] = {20, 20}, db[1] = {8};
      fmcs_entry e[2] = {{0, 0, 1, 20}, {0, 1, 1, 20}};
      void* dst[1] = {q->vals2};
      return fmcs_stitch(c, q->n, 2, src, sb, sw, 2, e, 1, dst, db);
    }, &a);
    out->stitch_ns_per_byte = t_s * 1e6 / (n * 16.0);
  }
  return FMCS_OK;
}

}  // extern "C"
