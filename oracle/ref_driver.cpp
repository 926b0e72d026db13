// ref_driver.cpp -- a thin C entry layer over the reference's own sources,
// compiled (by oracle/Makefile, target `ref`) from /root/reference/src where
// they lie.  It is TEST INFRASTRUCTURE: it exists to cross-check the oracle
// and to time the reference's AVX2 path as the CPU baseline.  Nothing in the
// product links it.
//
// The reference's round loop (src/composer.cpp:48-135) frees its results, so
// these wrappers call its per-round members directly and keep the outputs.
#include "composer.cpp"        // Composer::sortByGroup / subsort_mergesort / hashByGroup_* templates
#include "chain_composer.cpp"  // ChainComposer (concrete Composer)
#include <functional>          // standard headers generator.h uses without including them
#include <limits>
#include <random>
#include "generator.h"         // DataGenerator::ReadFromFile / ReadFromFile_tpch
#include "stitch_composer.cpp" // StitchComposer's estimators, greedy plans, cardinality statistics
#include "plan_enum_rrs.cpp"   // PlanEnumRRS (RRS over massaging plans)
#include <ctime>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <cstdint>
#include <thread>
#include <vector>

using namespace multiAttrSort;

namespace {

template <class T>
T *aligned_copy(const void *src, uint64_t n) {
  T *p = (T *)malloc_aligned((n + 64) * sizeof(T));
  std::memcpy(p, src, n * sizeof(T));
  return p;
}

Column *make_column(uint32_t bank, void *data) {
  switch (bank) {
    case 2: { auto *c = new SpecialiedColumn<uint16_t>(); c->SetColumn((uint16_t *)data); c->SetWidth(2); return c; }
    case 4: { auto *c = new SpecialiedColumn<uint32_t>(); c->SetColumn((uint32_t *)data); c->SetWidth(4); return c; }
    default: { auto *c = new SpecialiedColumn<uint64_t>(); c->SetColumn((uint64_t *)data); c->SetWidth(8); return c; }
  }
}

void *copy_bank(uint32_t bank, const void *src, uint64_t n) {
  switch (bank) {
    case 1: return aligned_copy<uint8_t>(src, n);
    case 2: return aligned_copy<uint16_t>(src, n);
    case 4: return aligned_copy<uint32_t>(src, n);
    default: return aligned_copy<uint64_t>(src, n);
  }
}

// ---- the round loop on several threads, over the reference's own kernels ----

template <class F>
void par(uint32_t nt, F f) {
  std::vector<std::thread> th;
  for (uint32_t t = 0; t < nt; ++t) th.emplace_back(f, t);
  for (auto &x : th) x.join();
}

// Rows [start, start + m) of bat sorted by their key: a stable radix
// partition on the key's top bits (per-thread histograms, the partitioning
// phase of the reference's pthread sort, src/mergesort.h:2603-2779), then
// every part sorted by the reference's subsort (avxmergesort_anytype), the
// parts spread over the threads.
template <class T>
void psort_range(ChainComposer &comp, BAT_t<T> *bat, uint64_t start, uint64_t m, uint32_t bitwidth, uint32_t nt,
                 T *tv, surrogate_t *to) {
  const uint32_t pb = bitwidth < 8 ? bitwidth : 8;
  const uint32_t P = 1u << pb;
  const uint32_t shift = bitwidth - pb;
  T *v = bat->values + start;
  surrogate_t *o = bat->oids + start;
  std::vector<uint64_t> hist((size_t)nt * P, 0), off((size_t)nt * P, 0), pstart(P + 1, 0);
  par(nt, [&](uint32_t t) {
    const uint64_t a = m * t / nt, b = m * (t + 1) / nt;
    uint64_t *h = hist.data() + (size_t)t * P;
    for (uint64_t i = a; i < b; ++i) ++h[((uint64_t)v[i] >> shift) & (P - 1)];
  });
  uint64_t run = 0;
  for (uint32_t p = 0; p < P; ++p) {
    pstart[p] = run;
    for (uint32_t t = 0; t < nt; ++t) {
      off[(size_t)t * P + p] = run;
      run += hist[(size_t)t * P + p];
    }
  }
  pstart[P] = m;
  par(nt, [&](uint32_t t) {
    const uint64_t a = m * t / nt, b = m * (t + 1) / nt;
    uint64_t *ofs = off.data() + (size_t)t * P;
    for (uint64_t i = a; i < b; ++i) {
      const uint64_t d = ofs[((uint64_t)v[i] >> shift) & (P - 1)]++;
      tv[d] = v[i];
      to[d] = o[i];
    }
  });
  par(nt, [&](uint32_t t) {
    const uint64_t a = m * t / nt, b = m * (t + 1) / nt;
    std::memcpy(v + a, tv + a, (b - a) * sizeof(T));
    std::memcpy(o + a, to + a, (b - a) * sizeof(surrogate_t));
  });
  // largest parts first, so a heavy part does not start last
  std::vector<uint32_t> ord(P);
  for (uint32_t p = 0; p < P; ++p) ord[p] = p;
  std::sort(ord.begin(), ord.end(),
            [&](uint32_t x, uint32_t y) { return pstart[x + 1] - pstart[x] > pstart[y + 1] - pstart[y]; });
  std::atomic<uint32_t> next(0);
  par(nt, [&](uint32_t) {
    for (uint32_t j; (j = next.fetch_add(1)) < P;) {
      const uint32_t p = ord[j];
      const uint64_t np = pstart[p + 1] - pstart[p];
      if (np > 1) comp.subsort<T>(bat, start + pstart[p], np, true, bitwidth);
    }
  });
}

// One round of Composer::sortByGroup<T> (src/composer.cpp:1222-1408) on nt
// threads: BATproject, the per-group sorts (groups above n / (8 nt) rows
// partitioned as above, the others spread over the threads in chunks), and
// extractGroupInfo (a two-pass parallel scan) unless it is the last round.
template <class T>
void round_mt(ChainComposer &comp, uint32_t nt, const T *col, uint64_t n, surrogate_t *order, surrogate_t *inGroup,
              surrogate_t *outGroup, uint32_t bitwidth, bool last) {
  T *vals = (T *)malloc_aligned(n * sizeof(T));
  T *tv = (T *)malloc_aligned(n * sizeof(T));
  surrogate_t *to = (surrogate_t *)malloc_aligned(n * sizeof(surrogate_t));
  if (inGroup)
    par(nt, [&](uint32_t t) {
      const uint64_t a = n * t / nt, b = n * (t + 1) / nt;
      for (uint64_t i = a; i < b; ++i) vals[i] = col[order[i]];
    });
  else
    std::memcpy(vals, col, n * sizeof(T));
  BAT_t<T> bat{order, vals, n};
  if (!inGroup) {
    psort_range<T>(comp, &bat, 0, n, bitwidth, nt, tv, to);
  } else {
    // group starts, found per thread chunk, then concatenated in order
    std::vector<std::vector<uint64_t>> st(nt);
    par(nt, [&](uint32_t t) {
      const uint64_t a = n * t / nt, b = n * (t + 1) / nt;
      for (uint64_t i = a; i < b; ++i)
        if (i == 0 || inGroup[i] != inGroup[i - 1]) st[t].push_back(i);
    });
    std::vector<uint64_t> starts;
    for (auto &x : st) starts.insert(starts.end(), x.begin(), x.end());
    starts.push_back(n);
    const uint64_t big = std::max<uint64_t>(4096, n / (8 * (uint64_t)nt));
    std::vector<uint64_t> small;
    for (size_t g = 0; g + 1 < starts.size(); ++g) {
      const uint64_t m = starts[g + 1] - starts[g];
      if (m >= big)
        psort_range<T>(comp, &bat, starts[g], m, bitwidth, nt, tv + starts[g], to + starts[g]);
      else if (m > 1)
        small.push_back(g);
    }
    std::atomic<size_t> next(0);
    const size_t CH = 64;
    par(nt, [&](uint32_t) {
      for (size_t c; (c = next.fetch_add(CH)) < small.size();)
        for (size_t j = c; j < std::min(small.size(), c + CH); ++j) {
          const uint64_t g = small[j];
          comp.subsort<T>(&bat, starts[g], starts[g + 1] - starts[g], true, bitwidth);
        }
    });
  }
  if (!last) {
    std::vector<uint64_t> heads(nt + 1, 0);
    auto head = [&](uint64_t i) -> uint32_t {
      return i > 0 && (vals[i] != vals[i - 1] || (inGroup && inGroup[i] != inGroup[i - 1]));
    };
    par(nt, [&](uint32_t t) {
      const uint64_t a = n * t / nt, b = n * (t + 1) / nt;
      uint64_t c = 0;
      for (uint64_t i = a; i < b; ++i) c += head(i);
      heads[t + 1] = c;
    });
    for (uint32_t t = 0; t < nt; ++t) heads[t + 1] += heads[t];
    par(nt, [&](uint32_t t) {
      const uint64_t a = n * t / nt, b = n * (t + 1) / nt;
      uint64_t g = heads[t];
      for (uint64_t i = a; i < b; ++i) {
        g += head(i);
        outGroup[i] = (surrogate_t)g;
      }
    });
  }
  free(vals);
  free(tv);
  free(to);
}

}  // namespace

extern "C" {

// Composer::subsort_mergesort<T> on one whole array (src/composer.cpp:490-556).
int ref_mergesort(uint32_t bank, void *keys, uint32_t *oids, uint64_t n) {
  setting_t s;
  compose_params_t p;
  ChainComposer comp(s, 1, n, p, nullptr);
  void *k = copy_bank(bank, keys, n);
  uint32_t *o = aligned_copy<uint32_t>(oids, n);
  switch (bank) {
    case 2: { BAT_t<uint16_t> b{o, (uint16_t *)k, n}; comp.subsort<uint16_t>(&b, 0, n, true, 16); break; }
    case 4: { BAT_t<uint32_t> b{o, (uint32_t *)k, n}; comp.subsort<uint32_t>(&b, 0, n, true, 32); break; }
    case 8: { BAT_t<uint64_t> b{o, (uint64_t *)k, n}; comp.subsort<uint64_t>(&b, 0, n, true, 64); break; }
    default: free(k); free(o); return -1;
  }
  std::memcpy(keys, k, n * bank);
  std::memcpy(oids, o, n * 4);
  free(k);
  free(o);
  return 0;
}

// The round loop of Composer::multiRoundSorting (src/composer.cpp:48-135)
// with sortByGroup<T> (:1222-1408), keeping order and per-round groups.
int ref_sort_rounds(uint32_t nrounds, const void *const *cols, const uint32_t *banks,
                    const uint32_t *bitwidths, uint64_t n, uint32_t *order_out,
                    uint32_t *groups_out) {
  setting_t s;
  compose_params_t p;
  bool asc[64];
  uint32_t bws[64];
  Column *colv[64];
  void *data[64];
  if (nrounds == 0 || nrounds > 64) return -1;
  for (uint32_t r = 0; r < nrounds; ++r) {
    asc[r] = true;
    bws[r] = bitwidths[r];
    data[r] = copy_bank(banks[r], cols[r], n);
    colv[r] = make_column(banks[r], data[r]);
  }
  p.column_asc_desc = asc;
  p.bitwidth = bws;
  ChainComposer comp(s, nrounds, n, p, colv);
  surrogate_t *order = (surrogate_t *)malloc_aligned(n * sizeof(surrogate_t));
  for (uint64_t i = 0; i < n; ++i) order[i] = (surrogate_t)i;
  surrogate_t *outGroup = (surrogate_t *)malloc_aligned(n * sizeof(surrogate_t));
  surrogate_t *inGroup = nullptr;
  for (uint32_t r = 0; r < nrounds; ++r) {
    bool last = (r == nrounds - 1);
    switch (banks[r]) {
      case 2: comp.sortByGroup<uint16_t>(outGroup, (uint16_t *)data[r], order, inGroup, true, bws[r], last); break;
      case 4: comp.sortByGroup<uint32_t>(outGroup, (uint32_t *)data[r], order, inGroup, true, bws[r], last); break;
      default: comp.sortByGroup<uint64_t>(outGroup, (uint64_t *)data[r], order, inGroup, true, bws[r], last); break;
    }
    if (!last && groups_out) std::memcpy(groups_out + (uint64_t)r * n, outGroup, n * sizeof(surrogate_t));
    if (!inGroup) inGroup = (surrogate_t *)malloc_aligned(n * sizeof(surrogate_t));
    swap<surrogate_t>(&inGroup, &outGroup);
  }
  std::memcpy(order_out, order, n * sizeof(surrogate_t));
  free(order);
  free(inGroup);
  free(outGroup);
  for (uint32_t r = 0; r < nrounds; ++r) { free(data[r]); delete colv[r]; }
  return 0;
}

// The round loop of ref_sort_rounds on nthreads threads (the reference's
// kernels, a parallel driver written here: its active path is single
// threaded, and its pthread sort, src/mergesort.h:290-441, has no caller).
// Orders agree with ref_sort_rounds up to the order of equal keys.
int ref_sort_rounds_mt(uint32_t nthreads, uint32_t nrounds, const void *const *cols, const uint32_t *banks,
                       const uint32_t *bitwidths, uint64_t n, uint32_t *order_out, uint32_t *groups_out) {
  if (nrounds == 0 || nrounds > 64 || nthreads == 0 || n == 0) return -1;
  setting_t s;
  compose_params_t p;
  bool asc[64];
  uint32_t bws[64];
  for (uint32_t r = 0; r < nrounds; ++r) {
    asc[r] = true;
    bws[r] = bitwidths[r];
  }
  p.column_asc_desc = asc;
  p.bitwidth = bws;
  ChainComposer comp(s, nrounds, n, p, nullptr);
  surrogate_t *order = (surrogate_t *)malloc_aligned(n * sizeof(surrogate_t));
  par(nthreads, [&](uint32_t t) {
    const uint64_t a = n * t / nthreads, b = n * (t + 1) / nthreads;
    for (uint64_t i = a; i < b; ++i) order[i] = (surrogate_t)i;
  });
  surrogate_t *outGroup = (surrogate_t *)malloc_aligned(n * sizeof(surrogate_t));
  surrogate_t *inGroup = nullptr;
  for (uint32_t r = 0; r < nrounds; ++r) {
    const bool last = (r == nrounds - 1);
    switch (banks[r]) {
      case 2: round_mt<uint16_t>(comp, nthreads, (const uint16_t *)cols[r], n, order, inGroup, outGroup, bws[r], last); break;
      case 4: round_mt<uint32_t>(comp, nthreads, (const uint32_t *)cols[r], n, order, inGroup, outGroup, bws[r], last); break;
      default: round_mt<uint64_t>(comp, nthreads, (const uint64_t *)cols[r], n, order, inGroup, outGroup, bws[r], last); break;
    }
    if (!last && groups_out) std::memcpy(groups_out + (uint64_t)r * n, outGroup, n * sizeof(surrogate_t));
    if (!inGroup) inGroup = (surrogate_t *)malloc_aligned(n * sizeof(surrogate_t));
    swap<surrogate_t>(&inGroup, &outGroup);
  }
  std::memcpy(order_out, order, n * sizeof(surrogate_t));
  free(order);
  free(inGroup);
  free(outGroup);
  return 0;
}

// ShiftLeftBatch + AssembleFromColumn over a construct-entry list
// (src/stitch_composer.cpp:652-703, src/column.h:105-1059).
int ref_stitch(uint64_t n, const void *const *src, const uint32_t *src_bank, const uint32_t *src_bw,
               uint32_t nent, const uint32_t *ent_newcol, const uint32_t *ent_src,
               const uint32_t *ent_start, const uint32_t *ent_end, uint32_t ndst, void *const *dst,
               const uint32_t *dst_bank) {
  Column *srcc[64];
  void *srcd[64];
  uint32_t nsrc = 0;
  for (uint32_t e = 0; e < nent; ++e) nsrc = ent_src[e] + 1 > nsrc ? ent_src[e] + 1 : nsrc;
  if (nsrc > 64 || ndst > 64) return -1;
  for (uint32_t s = 0; s < nsrc; ++s) {
    srcd[s] = copy_bank(src_bank[s], src[s], n);
    srcc[s] = make_column(src_bank[s], srcd[s]);
  }
  for (uint32_t c = 0; c < ndst; ++c) {
    void *d = copy_bank(dst_bank[c], dst[c], n);
    std::memset(d, 0, n * dst_bank[c]);
    Column *dc = make_column(dst_bank[c], d);
    bool first = true;
    for (uint32_t e = 0; e < nent; ++e) {
      if (ent_newcol[e] != c) continue;
      uint32_t w = ent_end[e] - ent_start[e] + 1;
      if (!first) dc->ShiftLeftBatch(n, w);
      first = false;
      uint32_t s = ent_src[e];
      switch (src_bank[s]) {
        case 2: dc->AssembleFromColumn(reinterpret_cast<SpecialiedColumn<uint16_t> *>(srcc[s]), ent_start[e], ent_end[e], n, src_bw[s]); break;
        case 4: dc->AssembleFromColumn(reinterpret_cast<SpecialiedColumn<uint32_t> *>(srcc[s]), ent_start[e], ent_end[e], n, src_bw[s]); break;
        default: dc->AssembleFromColumn(reinterpret_cast<SpecialiedColumn<uint64_t> *>(srcc[s]), ent_start[e], ent_end[e], n, src_bw[s]); break;
      }
    }
    std::memcpy(dst[c], d, n * dst_bank[c]);
    free(d);
    delete dc;
  }
  for (uint32_t s = 0; s < nsrc; ++s) { free(srcd[s]); delete srcc[s]; }
  return 0;
}

// DataGenerator::ReadFromFile (tpch == 0) / ReadFromFile_tpch
// (src/generator.h:67-136): copies column c's values into cols[c] (its
// bank: <= 16 bits u16, <= 32 u32, else u64) and its std::set cardinality
// into card[c]; *nrows is the caller's count, or gets the TPC-H header's.
// The reference exits on a missing file, so the caller passes existing
// files only.
int ref_read_text(const char *path, int tpch, uint32_t ncols, const uint32_t *widths, uint64_t *nrows,
                  void *const *cols, uint64_t *card) {
  if (ncols > 64) return -1;
  Column *colv[64];
  uint32_t w[64];
  for (uint32_t c = 0; c < ncols; ++c) w[c] = widths[c];
  DataGenerator g;
  uint64_t n = *nrows;
  if (tpch)
    g.ReadFromFile_tpch(colv, std::string(path), ncols, n, w);
  else
    g.ReadFromFile(colv, std::string(path), ncols, n, w);
  *nrows = n;
  for (uint32_t c = 0; c < ncols; ++c) {
    if (cols[c]) std::memcpy(cols[c], colv[c]->GetColumn(), n * colv[c]->GetWidth());
    card[c] = colv[c]->GetCardinality();
    free(colv[c]->GetColumn());
    delete colv[c];
  }
  return 0;
}

}  // extern "C"

// ---- the reference's planner ------------------------------------------------
//
// Its statistics, group-count estimate, column cost, greedy plans and plan
// costing are private members of StitchComposer / PlanEnumRRS.  A pointer to
// a private member may be named in an explicit instantiation (access is not
// checked there), and a friend defined by that instantiation hands it out.
namespace {

template <class Tag, typename Tag::type M>
struct Reach {
  friend typename Tag::type reach(Tag) { return M; }
};

struct CardStats {
  typedef void (StitchComposer::*type)(uint32_t *, std::vector<uint32_t> &, double *, uint32_t (*)[3], uint32_t *);
  friend type reach(CardStats);
};
template struct Reach<CardStats, &StitchComposer::getCardinalityFromStatistics>;

struct GroupNum {
  typedef double (StitchComposer::*type)(uint32_t, uint32_t *);
  friend type reach(GroupNum);
};
template struct Reach<GroupNum, &StitchComposer::estimateGroupNum>;

struct ColumnCost {
  typedef double (StitchComposer::*type)(double, uint32_t);
  friend type reach(ColumnCost);
};
template struct Reach<ColumnCost, &StitchComposer::calculate_single_column_cost>;

struct Greedy {
  typedef void (StitchComposer::*type)(uint32_t, std::vector<std::vector<uint32_t> > &, uint32_t *);
  friend type reach(Greedy);
};
template struct Reach<Greedy, &StitchComposer::find_greedy_plans>;

struct PlanCost {
  typedef void (StitchComposer::*type)(uint32_t *, std::vector<std::vector<uint32_t> > &, int &, double &);
  friend type reach(PlanCost);
};
template struct Reach<PlanCost, &StitchComposer::estimate_plan_collection_greedy>;

struct RrsStats {
  typedef void (PlanEnumRRS::*type)(const std::vector<uint32_t> &, uint64_t *, uint32_t (*)[3], uint32_t *);
  friend type reach(RrsStats);
};
template struct Reach<RrsStats, &PlanEnumRRS::GetStatistics>;

// A StitchComposer over columns that carry only their cardinality (the
// planner reads no values).
struct RefPlanner {
  std::vector<uint32_t> bw;
  std::vector<SpecialiedColumn<uint64_t> *> cols;
  std::vector<Column *> colp;
  StitchComposer *sc = nullptr;
  RefPlanner(uint64_t n, uint32_t ncols, const uint32_t *bws, const uint64_t *card) : bw(bws, bws + ncols) {
    uint32_t S = 0;
    for (uint32_t c = 0; c < ncols; ++c) {
      auto *col = new SpecialiedColumn<uint64_t>();
      col->SetColumn(nullptr);
      col->SetWidth(8);
      col->SetCardinality(card[c]);
      cols.push_back(col);
      colp.push_back(col);
      S += bws[c];
    }
    compose_params_t p;
    p.bitwidth = bw.data();
    p.colwidth_sum = S;
    setting_t st;
    sc = new StitchComposer(st, ncols, n, p, colp.data());
  }
  ~RefPlanner() {
    delete sc;
    for (auto *c : cols) delete c;
  }
};

struct RefRrs : PlanEnumRRS {
  using PlanEnumRRS::PlanEnumRRS;
  using PlanEnumRRS::GetCost;
};

std::vector<uint32_t> vec(const uint32_t *a, uint32_t n) { return std::vector<uint32_t>(a, a + n); }

int put_plans(const std::vector<std::vector<uint32_t> > &plans, uint32_t *out, uint64_t cap, uint64_t *nplans,
              uint64_t *words) {
  uint64_t w = 0;
  for (const auto &p : plans) {
    if (w + 1 + p.size() > cap) return -1;
    out[w++] = (uint32_t)p.size();
    for (uint32_t x : p) out[w++] = x;
  }
  *nplans = plans.size();
  *words = w;
  return 0;
}

}  // namespace

extern "C" {

void *ref_planner_new(uint64_t n, uint32_t ncols, const uint32_t *bw, const uint64_t *card) {
  return new RefPlanner(n, ncols, bw, card);
}
void ref_planner_free(void *h) { delete (RefPlanner *)h; }

// getCardinalityFromStatistics (src/stitch_composer.cpp:288-553)
void ref_card_stats(void *h, const uint32_t *order, uint32_t nsplit, const uint32_t *split, double *card,
                    uint32_t *assemble9, uint32_t *shift3) {
  RefPlanner *r = (RefPlanner *)h;
  std::vector<uint32_t> ord(order, order + r->bw.size()), sp = vec(split, nsplit);
  uint32_t a[3][3] = {{0}};
  uint32_t sh[3] = {0, 0, 0};
  for (uint32_t i = 0; i < nsplit; ++i) card[i] = 1;
  (r->sc->*reach(CardStats()))(ord.data(), sp, card, a, sh);
  std::memcpy(assemble9, a, sizeof(a));
  std::memcpy(shift3, sh, sizeof(sh));
}

// estimateGroupNum (src/stitch_composer.cpp:1845-1867)
double ref_group_num(void *h, uint32_t bitpos, const uint32_t *order) {
  RefPlanner *r = (RefPlanner *)h;
  std::vector<uint32_t> ord(order, order + r->bw.size());
  return (r->sc->*reach(GroupNum()))(bitpos, ord.data());
}

// calculate_single_column_cost (src/stitch_composer.cpp:2186-2226)
double ref_column_cost(void *h, double g, uint32_t bank_type) {
  return (((RefPlanner *)h)->sc->*reach(ColumnCost()))(g, bank_type);
}

// find_greedy_plans (src/stitch_composer.cpp:1992-2125), rounds > 2
int ref_greedy_plans(void *h, uint32_t rounds, const uint32_t *order, uint32_t *out, uint64_t cap, uint64_t *nplans,
                     uint64_t *words) {
  RefPlanner *r = (RefPlanner *)h;
  std::vector<uint32_t> ord(order, order + r->bw.size());
  std::vector<std::vector<uint32_t> > plans;
  (r->sc->*reach(Greedy()))(rounds, plans, ord.data());
  return put_plans(plans, out, cap, nplans, words);
}

// One plan through estimate_plan_collection_greedy (src/stitch_composer.cpp:2228-2378)
double ref_plan_cost(void *h, const uint32_t *order, uint32_t nsplit, const uint32_t *split) {
  RefPlanner *r = (RefPlanner *)h;
  std::vector<uint32_t> ord(order, order + r->bw.size());
  std::vector<std::vector<uint32_t> > plans{vec(split, nsplit)};
  int best = -1;
  double cost = 0;
  (r->sc->*reach(PlanCost()))(ord.data(), plans, best, cost);
  return cost;
}

// PlanEnumRRS::GetStatistics and GetCost (src/plan_enum_rrs.cpp:94-197, :269-396)
double ref_rrs_cost(uint64_t n, uint32_t max_round, const uint32_t *order, uint32_t ncols, const uint32_t *bw,
                    const uint64_t *card, uint32_t nsplit, const uint32_t *split, uint64_t *card_out,
                    uint32_t *assemble9, uint32_t *shift3) {
  uint32_t S = 0;
  std::vector<uint32_t> c32(ncols);
  for (uint32_t c = 0; c < ncols; ++c) {
    S += bw[c];
    c32[c] = (uint32_t)card[c];
  }
  RefRrs e(n, S, max_round, vec(order, ncols), ncols, vec(bw, ncols), c32, 1e18);
  std::vector<uint32_t> sp = vec(split, nsplit);
  if (card_out) {
    uint32_t a[3][3] = {{0}};
    uint32_t sh[3] = {0, 0, 0};
    for (uint32_t i = 0; i < nsplit; ++i) card_out[i] = 1;
    (e.*reach(RrsStats()))(sp, card_out, a, sh);
    std::memcpy(assemble9, a, sizeof(a));
    std::memcpy(shift3, sh, sizeof(sh));
  }
  return e.GetCost(sp);
}

// PlanEnumRRS::Optimize and ObtainPlans (src/plan_enum_rrs.h:20-79): its engine
// is seeded from time(0) in the constructor (src/rrs.h:68); the second read
// before and after the construction tells the seed (retried when they differ).
int ref_plan_enum_rrs(uint64_t n, uint32_t max_round, const uint32_t *order, uint32_t ncols, const uint32_t *bw,
                      const uint64_t *card, uint64_t *seed, uint32_t *out, uint64_t cap, uint64_t *nplans,
                      uint64_t *words) {
  uint32_t S = 0;
  std::vector<uint32_t> c32(ncols);
  for (uint32_t c = 0; c < ncols; ++c) {
    S += bw[c];
    c32[c] = (uint32_t)card[c];
  }
  for (;;) {
    const std::time_t t0 = std::time(0);
    RefRrs e(n, S, max_round, vec(order, ncols), ncols, vec(bw, ncols), c32, 1e18);
    const std::time_t t1 = std::time(0);
    if (t0 != t1) continue;
    e.Optimize();
    const auto visited = e.ObtainPlans();
    *seed = (uint64_t)t0;
    return put_plans(std::vector<std::vector<uint32_t> >(visited.begin(), visited.end()), out, cap, nplans, words);
  }
}

// The reference's static estimators with the signatures of the product's
// fmcs_cost_primitives callbacks (src/stitch_composer.h:54-72).
double ref_cb_sort(void *, double g, uint32_t bank_type) {
  switch (bank_type) {
    case 0: return StitchComposer::estimate_mergesort_uint16(g);
    case 1: return StitchComposer::estimate_mergesort_uint32(g);
    default: return StitchComposer::estimate_mergesort_uint64(g);
  }
}
double ref_cb_stitch(void *, const uint32_t *assemble9, const uint32_t *shift3, uint64_t n) {
  uint32_t a[3][3], sh[3];
  std::memcpy(a, assemble9, sizeof(a));
  std::memcpy(sh, shift3, sizeof(sh));
  return StitchComposer::estimate_stitch(a, sh, n);  // nitems is uint32_t there, as in its callers
}
double ref_cb_reconstruct(void *, const uint32_t *bytewidths, uint32_t nrounds, uint64_t n) {
  return StitchComposer::estimate_tuple_reconstruct(bytewidths, nrounds, n);
}

}  // extern "C"
