// fmcs_plan.cpp -- host-side massaging planner: plan enumeration, the
// analytic cost model (structure of the reference's, constants measured on
// the GPU), greedy plans, Recursive Random Search and the optimizer.
// Pure host C++; no GPU is touched here.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <random>
#include <functional>
#include <set>
#include <type_traits>
#include <vector>

#include "../../include/fmcs.h"
#include "../../include/fmcs_plan.h"
#include "fmcs_rrs.h"

namespace {

typedef std::vector<uint32_t> Split;

int bank_type(uint32_t width) { return width <= 16 ? 0 : (width <= 32 ? 1 : 2); }

// ---- cost model ---------------------------------------------------------

// Per-row segmented-sort time for an average group size g (log2 buckets,
// linear interpolation in log2 g).
double segsort_ns(const fmcs_cost_model* m, int bt, double g) {
  double lg = std::log2(std::max(1.0, g));
  int i = (int)std::floor(lg);
  if (i >= FMCS_COST_GBUCKETS - 1) return m->segsort_ns[bt][FMCS_COST_GBUCKETS - 1];
  double f = lg - i;
  return m->segsort_ns[bt][i] * (1 - f) + m->segsort_ns[bt][i + 1] * f;
}

// The round-1 radix sort cost: one LSD pass per 8 key bits.
double radix_ms(const fmcs_cost_model* m, int bt, uint32_t bits, double n) {
  uint32_t passes = (bits + 7) / 8;
  return n * passes * m->radix_pass_ns[bt] * 1e-6;
}

// The bank type (0: u16, 1: u32, 2: u64) of a new column's width.
// calculate_single_column_cost (src/stitch_composer.cpp:2186-2226,
// src/plan_enum_rrs.cpp:7-35): (n/g) groups times the cost of sorting one
// group of g rows, scaled by f(g) = 1/(1 + 1/(g-1)) (a group of one row
// needs no sort).  `one_group` prices one group.
template <class F>
double column_cost(double n, double g, F one_group) {
  const double group_num = n / g;
  const double factor = (g - 1.0) < 0.000001 ? 0.0 : 1.0 / (1.0 + 1.0 / (g - 1.0));
  return group_num * factor * one_group(g);
}

// The same with per-row GPU times: one group of g rows costs g * per_row(g).
double single_column_cost(const fmcs_cost_model* m, double n, double g, int bt) {
  return column_cost(n, g, [&](double x) { return x * segsort_ns(m, bt, x) * 1e-6; });
}

// ... and with caller-supplied primitives (the reference's sort estimators).
double single_column_cost_prim(const fmcs_cost_primitives* p, double n, double g, int bt) {
  return column_cost(n, g, [&](double x) { return p->sort_cost(p->user, x, (uint32_t)bt); });
}

struct Stats {
  uint32_t assemble[3][3];  // [from bank][to bank] entry counts
  uint32_t shift[3];        // entries per destination bank
};

// The conversion of a double to uint64 that the reference's cardinalities
// go through (`uint64 *= double`, src/plan_enum_rrs.cpp:393) on x86-64:
// truncation below 2^63, exact in [2^63, 2^64), 0 from 2^64 up and 2^63
// for NaN (cvttsd2si with its out-of-range value, and the 2^63 offset).
uint64_t u64_of(double v) {
  if (std::isnan(v)) return 1ull << 63;
  if (v >= 18446744073709551616.0) return 0;
  if (v >= 9223372036854775808.0) return (uint64_t)(int64_t)(v - 9223372036854775808.0) ^ (1ull << 63);
  if (v < -9223372036854775808.0) return 1ull << 63;
  return (uint64_t)(int64_t)v;
}

// getCardinalityFromStatistics (src/stitch_composer.cpp:288-553) and
// PlanEnumRRS::GetStatistics (src/plan_enum_rrs.cpp:269-396): walk the
// construct entries of the split, count assemble/shift operations by bank,
// and multiply card_per_bit^width into each new column's cardinality (all
// but the last).  Card is double (the optimizer's variant) or uint64_t (the
// RRS variant, converted after every entry).
template <class Card>
void plan_stats(uint32_t ncols, const uint32_t* bw, const uint64_t* card, const uint32_t* order, const Split& split,
                std::vector<Card>& cardinality, Stats& st) {
  std::memset(&st, 0, sizeof(st));
  const size_t nn = split.size();
  cardinality.assign(nn, (Card)1);
  std::vector<double> cpb(ncols);
  for (uint32_t c = 0; c < ncols; ++c) cpb[c] = std::pow((double)card[c], 1.0 / (double)bw[c]);
  uint32_t oi = 0, ni = 0;
  uint32_t ow = bw[order[0]], nw = split[0];
  while (ni < nn) {
    uint32_t m = std::min(ow, nw);
    ow -= m;
    nw -= m;
    uint32_t src = order[oi];
    int to = bank_type(split[ni]);
    int from = bank_type(bw[src]);
    st.shift[to]++;
    st.assemble[from][to]++;
    if (ni + 1 < nn) {
      const double v = (double)cardinality[ni] * std::pow(cpb[src], (double)m);
      if constexpr (std::is_same<Card, uint64_t>::value)
        cardinality[ni] = u64_of(v);
      else
        cardinality[ni] = v;
    }
    if (ow == 0) {
      ++oi;
      if (oi < ncols) ow = bw[order[oi]];
    }
    if (nw == 0) {
      ++ni;
      if (ni < nn) nw = split[ni];
    }
  }
}

// Groups before each round from the new columns' cardinalities: G_0 = 1,
// G_r = card_0 * ... * card_{r-1} (src/stitch_composer.cpp:2306-2318; in
// uint64 with wrap-around in the RRS variant, src/plan_enum_rrs.cpp:149-156).
template <class Card>
std::vector<double> round_groups(const std::vector<Card>& cardinality) {
  std::vector<double> G(cardinality.size());
  Card acc = cardinality.empty() ? (Card)1 : cardinality[0];
  if (!G.empty()) G[0] = 1.0;
  for (size_t r = 1; r < cardinality.size(); ++r) {
    G[r] = (double)acc;
    acc *= cardinality[r];
  }
  return G;
}

std::vector<double> plan_groups(uint32_t ncols, const uint32_t* bw, const uint64_t* card, const uint32_t* order,
                                const Split& split, bool truncate, Stats& st) {
  if (truncate) {
    std::vector<uint64_t> cu;
    plan_stats(ncols, bw, card, order, split, cu, st);
    return round_groups(cu);
  }
  std::vector<double> cd;
  plan_stats(ncols, bw, card, order, split, cd, st);
  return round_groups(cd);
}

// estimate_plan_collection_greedy's per-plan costing (src/stitch_composer.cpp:
// 2228-2378) and PlanEnumRRS::GetCost (src/plan_enum_rrs.cpp:94-197) with
// caller-supplied primitives: stitch + tuple reconstruction + (group
// extraction, 0 there) + the per-round sorts, in that order.
double estimate_prim(const fmcs_cost_primitives* p, double n, uint32_t ncols, const uint32_t* bw,
                     const uint64_t* card, const uint32_t* order, const Split& split, bool truncate) {
  Stats st;
  const std::vector<double> G = plan_groups(ncols, bw, card, order, split, truncate, st);
  const size_t R = split.size();
  std::vector<uint32_t> bytes(R);
  double sort = 0.0;
  for (size_t r = 0; r < R; ++r) {
    bytes[r] = fmcs_bank_bytes(split[r], 0);
    sort += single_column_cost_prim(p, n, n / G[r], bank_type(split[r]));
  }
  const double stitch = p->stitch_cost(p->user, &st.assemble[0][0], st.shift, (uint64_t)n);
  const double recon = p->reconstruct_cost(p->user, bytes.data(), (uint32_t)R, (uint64_t)n);
  const double extract = 0.0;
  return stitch + recon + extract + sort;
}

// The GPU model: stitch, tuple reconstruction (gather), multi-pass sort and
// group extraction, with the executor's fixed costs.
double estimate(const fmcs_cost_model* m, double n, uint32_t ncols, const uint32_t* bw, const uint64_t* card,
                const uint32_t* order, const Split& split, bool truncate, double* parts) {
  Stats st;
  const std::vector<double> G = plan_groups(ncols, bw, card, order, split, truncate, st);
  const size_t R = split.size();
  // A carried second column makes every round-1 pass move its bank too
  // (the passes are HBM-bound: time scales with the bytes moved per row,
  // key + oid + payload against key + oid) and removes round 2's gather.
  // Packed into u64 composites, a round-1 pass moves what a u64-bank pass
  // moves (composite + oid).
  const bool carry = m->carry_next && R >= 2 && n > 1;
  const bool fused = carry && m->carry_next == 1 && fmcs_bank_bytes(split[0], 0) <= 4 &&
                     fmcs_bank_bytes(split[1], 0) == 8 && split[0] + split[1] <= 64;
  double sort_ms = 0, gather_ms = 0, extract_ms = 0;
  for (size_t r = 0; r < R; ++r) {
    int bt = bank_type(split[r]);
    if (r == 0) {
      double scale = 1.0;
      if (carry && !fused)
        scale += (double)fmcs_bank_bytes(split[1], 0) / ((double)fmcs_bank_bytes(split[0], 0) + 4.0);
      sort_ms += fused ? radix_ms(m, 2, split[0], n) : radix_ms(m, bt, split[0], n) * scale;
      sort_ms += ((split[0] + 7) / 8) * m->pass_overhead_ms;
    } else {
      double g = std::max(1.0, n / std::min(G[r], n));
      sort_ms += single_column_cost(m, n, g, bt);
      // the device-driven level schedule: one level per 8 key bits and a last
      // one, unless no segment can exceed a tile (one level)
      const double levels = n <= 2048.0 ? 1.0 : (double)((split[r] + 7) / 8 + 1);
      sort_ms += levels * m->level_overhead_ms;
      if (!(carry && r == 1)) gather_ms += n * m->gather_ns[bt] * 1e-6;
    }
    if (r + 1 < R) extract_ms += n * m->boundary_ns[bt] * 1e-6;
  }
  // fused stitch: every source bank read once, every new bank written once
  double src_bytes = 0, dst_bytes = 0;
  for (uint32_t c = 0; c < ncols; ++c) src_bytes += fmcs_bank_bytes(bw[c], 0);
  for (size_t r = 0; r < R; ++r) dst_bytes += fmcs_bank_bytes(split[r], 0);
  double stitch_ms = n * (src_bytes + dst_bytes) * m->stitch_ns_per_byte * 1e-6;
  double total = stitch_ms + gather_ms + sort_ms + extract_ms + R * m->round_overhead_ms;
  if (parts) {
    parts[0] = stitch_ms;
    parts[1] = gather_ms;
    parts[2] = sort_ms;
    parts[3] = extract_ms;
  }
  return total;
}

// ---- enumerators -----------------------------------------------------------

// split_exhaustive_certain_round (src/stitch_composer.cpp:205-230)
void split_rec(int n, int cur, int rounds, int maxv, uint32_t* part, std::vector<Split>& out) {
  if (cur == rounds - 1) {
    if (n > maxv) return;
    part[cur] = n;
    out.emplace_back(part, part + rounds);
    return;
  }
  for (int i = n > maxv ? maxv : n; i > 0; --i) {
    part[cur] = i;
    if (n - i > 0) split_rec(n - i, cur + 1, rounds, maxv, part, out);
  }
}

std::vector<Split> split_exhaustive(uint32_t S, uint32_t rounds, uint32_t maxv) {
  std::vector<Split> out;
  if (rounds == 0 || rounds > 64 || S == 0) return out;
  uint32_t part[64];
  split_rec((int)S, 0, (int)rounds, (int)maxv, part, out);
  return out;
}

// enum_banksize_category (src/stitch_composer.cpp:1717-1734): 3^r patterns
// in lexicographic order.
std::vector<Split> bank_categories(uint32_t rounds) {
  std::vector<Split> out;
  Split cur(rounds, 0);
  uint64_t total = 1;
  for (uint32_t i = 0; i < rounds; ++i) total *= 3;
  for (uint64_t k = 0; k < total; ++k) {
    uint64_t x = k;
    for (int i = (int)rounds - 1; i >= 0; --i) {
      cur[i] = (uint32_t)(x % 3);
      x /= 3;
    }
    out.push_back(cur);
  }
  return out;
}

// estimateGroupNum (src/stitch_composer.cpp:1845-1867)
double estimate_group_num(uint32_t bit_position, uint32_t ncols, const uint32_t* bw, const uint64_t* card,
                          const uint32_t* order) {
  uint32_t avail = bit_position;
  double g = 1.0;
  for (uint32_t c = 0; c < ncols; ++c) {
    uint32_t src = order[c];
    uint32_t take = std::min(avail, bw[src]);
    g *= std::pow(std::pow((double)card[src], 1.0 / (double)bw[src]), (double)take);
    avail -= take;
    if (avail == 0) break;
  }
  return g;
}

// find_greedy_plans (src/stitch_composer.cpp:1992-2125): per bank pattern,
// every cut is placed where the NEXT round's sort is cheapest
// (determineBitRange :1736-1763, determineOptBitPosition_v2 :2155-2184; a
// tie goes to the later position).  colcost(n, g, bank type) prices a
// round over groups of g rows.
template <class ColCost>
std::vector<Split> greedy_plans(ColCost colcost, double n, uint32_t ncols, const uint32_t* bw, const uint64_t* card,
                                const uint32_t* order, uint32_t rounds) {
  std::vector<Split> out;
  uint32_t S = 0;
  for (uint32_t c = 0; c < ncols; ++c) S += bw[c];
  for (const Split& cate : bank_categories(rounds)) {
    Split plan;
    uint32_t off = 0;
    bool valid = true;
    for (uint32_t r = 0; r + 1 < rounds; ++r) {
      int lo = 0, hi = 0;
      switch (cate[r]) {
        case 0: lo = off + 1; hi = off + 16; break;
        case 1: lo = off + 17; hi = off + 32; break;
        default: lo = off + 33; hi = off + 64; break;
      }
      if (lo > (int)S) { valid = false; break; }
      hi = std::min(hi, (int)S);
      double best = std::numeric_limits<double>::max();
      int opt = -1;
      for (int pos = lo; pos <= hi; ++pos) {
        const double gsz = n / estimate_group_num((uint32_t)pos, ncols, bw, card, order);
        const double c = colcost(n, gsz, (int)cate[r + 1]);
        if (c <= best) {
          best = c;
          opt = pos;
        }
      }
      if (opt <= (int)off) { valid = false; break; }
      plan.push_back((uint32_t)opt - off);
      off = (uint32_t)opt;
    }
    if (!valid) continue;
    uint32_t last = S - off;
    switch (cate[rounds - 1]) {
      case 0: valid = last > 0 && last <= 16; break;
      case 1: valid = last > 16 && last <= 32; break;
      default: valid = last > 32 && last <= 64; break;
    }
    if (valid) {
      plan.push_back(last);
      out.push_back(plan);
    }
  }
  return out;
}

std::vector<Split> greedy_plans(const fmcs_cost_model* m, double n, uint32_t ncols, const uint32_t* bw,
                                const uint64_t* card, const uint32_t* order, uint32_t rounds) {
  auto cc = [m](double nn, double g, int bt) { return single_column_cost(m, nn, std::max(1.0, g), bt); };
  return greedy_plans(cc, n, ncols, bw, card, order, rounds);
}

// ---- Recursive Random Search over massaging plans ---------------------------
//
// PlanEnumRRS (src/plan_enum_rrs.cpp:37-267, parameters src/plan_enum_rrs.h:29):
// a plan is a split (round widths summing to S, each at most 64), costed with
// the RRS variant of the cardinalities (truncated to uint64).  The samples and
// neighbours follow the reference's draws one for one, so a seed visits the
// reference's plans: the sampler draws a round count once before its retry
// loop (src/plan_enum_rrs.cpp:48), and a neighbour move off either end of
// the split appends a new last round (:225-231).
class PlanEnumRRS : public fmcs::rrs::Search<Split, double> {
 public:
  PlanEnumRRS(std::function<double(const Split&)> cost, uint32_t S, uint32_t max_round, uint64_t seed)
      : Search<Split, double>(seed), cost_(std::move(cost)), S_(S), max_round_(max_round) {}

  const std::set<Split>& Plans() const { return plans_; }

 protected:
  // GetRandomSample (src/plan_enum_rrs.cpp:37-92)
  Split Sample() override {
    (void)Dice();  // the reference's first round-count draw, overwritten inside its loop
    Split s;
    for (;;) {
      s.clear();
      const uint32_t rounds = std::min<uint32_t>((uint32_t)std::ceil(Dice() * max_round_), max_round_);
      uint32_t avail = S_, cur = 0;
      while (cur + 1 < rounds) {
        uint32_t w = 1 + (uint32_t)std::floor(Dice() * (double)(avail - 1));
        w = std::min<uint32_t>(w, 64);
        s.push_back(w);
        avail -= w;
        ++cur;
        if (avail == 0) break;
      }
      if (avail > 0 && avail <= 64) {
        s.push_back(avail);
        plans_.insert(s);
        return s;
      }
    }
  }

  double Evaluate(const Split& s) override { return cost_(s); }

  // StopCritera (src/plan_enum_rrs.cpp:199-204)
  bool Stop() override { return ++counter_ > 10; }

  // GetRandomNeighbor (src/plan_enum_rrs.cpp:206-267): move single bits
  // between adjacent rounds, ceil(S * ro) moves
  Split Neighbor(const Split& pt, double ro) override {
    Split t(pt.begin(), pt.end());
    uint32_t steps = (uint32_t)std::ceil(S_ * ro);
    while (steps > 0) {
      int dec = (int)std::floor(Dice() * t.size());
      const uint32_t coin = ((uint32_t)(100.0 * Dice())) % 2;
      int inc = coin ? dec - 1 : dec + 1;
      if (inc == -1 || inc == (int)t.size()) {
        t.push_back(0);
        inc = (int)t.size() - 1;
      }
      if (t[dec] > 0 && t[inc] < 64) {
        t[dec]--;
        t[inc]++;
      }
      --steps;
    }
    Split out;
    for (uint32_t w : t)
      if (w) out.push_back(w);
    plans_.insert(out);
    return out;
  }

 private:
  std::function<double(const Split&)> cost_;
  uint32_t S_, max_round_;
  size_t counter_ = 0;
  std::set<Split> plans_;
};

int write_plans(const std::vector<Split>& plans, uint32_t* out, uint64_t cap_words, uint64_t* nplans,
                uint64_t* words) {
  uint64_t w = 0;
  for (const Split& s : plans) {
    if (out) {
      if (w + 1 + s.size() > cap_words) return FMCS_E_ARG;
      out[w] = (uint32_t)s.size();
      for (size_t i = 0; i < s.size(); ++i) out[w + 1 + i] = s[i];
    }
    w += 1 + s.size();
  }
  if (nplans) *nplans = plans.size();
  if (words) *words = w;
  return FMCS_OK;
}

bool inputs_ok(uint32_t ncols, const uint32_t* bw, const uint64_t* card) {
  if (ncols == 0 || ncols > 64 || !bw || !card) return false;
  for (uint32_t c = 0; c < ncols; ++c)
    if (bw[c] == 0 || bw[c] > 64 || card[c] == 0) return false;
  return true;
}

}  // namespace

extern "C" {

int fmcs_cost_model_default(fmcs_cost_model* m) {
  if (!m) return FMCS_E_ARG;
  std::memset(m, 0, sizeof(*m));
  // Constants measured on MI355X by fmcs_calibrate (2^24 rows, the
  // tuning file's settings); replaced by fmcs_calibrate() on the running
  // device.  Per-row ns: an LSD pass, a gather, a boundary pass, by bank
  // (u16, u32, u64); the segmented sort by bank and log2(average group size).
  static const double radix[3] = {0.007184, 0.00818, 0.009577};
  static const double gather[3] = {0.0163, 0.01823, 0.02006};
  static const double bnd[3] = {0.003929, 0.004761, 0.007405};
  static const double seg[3][FMCS_COST_GBUCKETS] = {
      {0.02143, 0.01415, 0.01041, 0.008636, 0.007965, 0.007262, 0.007022, 0.006923, 0.006961, 0.007016, 0.005465, 0.00553, 0.01742, 0.01667, 0.01532, 0.01396, 0.01325, 0.01446, 0.01622, 0.02776, 0.03676, 0.05208, 0.0504, 0.04944, 0.04838, 0.04838, 0.04838, 0.04838, 0.04838, 0.04838, 0.04838, 0.04838},
      {0.02094, 0.01374, 0.0105, 0.008983, 0.007981, 0.007363, 0.007074, 0.007217, 0.007168, 0.007068, 0.005336, 0.005388, 0.01783, 0.01717, 0.01566, 0.01416, 0.01343, 0.01403, 0.01384, 0.02194, 0.02403, 0.02391, 0.02299, 0.02188, 0.02101, 0.02101, 0.02101, 0.02101, 0.02101, 0.02101, 0.02101, 0.02101},
      {0.01966, 0.01238, 0.00922, 0.00761, 0.006579, 0.005799, 0.005619, 0.00561, 0.005563, 0.005657, 0.004405, 0.005248, 0.01788, 0.01912, 0.01688, 0.01523, 0.01461, 0.01537, 0.01574, 0.02443, 0.02603, 0.02656, 0.02477, 0.02375, 0.02302, 0.02302, 0.02302, 0.02302, 0.02302, 0.02302, 0.02302, 0.02302}};
  for (int b = 0; b < 3; ++b) {
    m->radix_pass_ns[b] = radix[b];
    m->gather_ns[b] = gather[b];
    m->boundary_ns[b] = bnd[b];
    for (int i = 0; i < FMCS_COST_GBUCKETS; ++i) m->segsort_ns[b][i] = seg[b][i];
  }
  m->stitch_ns_per_byte = 0.0001571;  // about 6.4 TB/s
  m->round_overhead_ms = 0.005;
  m->pass_overhead_ms = 0.02042;
  m->level_overhead_ms = 0.03186;
  m->carry_next = 1;
  return FMCS_OK;
}

int fmcs_estimate_plan(const fmcs_cost_model* m, uint64_t nrows, uint32_t ncols, const uint32_t* bw,
                       const uint64_t* card, const uint32_t* col_order, uint32_t nsplit, const uint32_t* split,
                       int truncate_cardinality, double* est_ms, double* parts4) {
  if (!m || !est_ms || !split || nsplit == 0 || !col_order || !inputs_ok(ncols, bw, card)) return FMCS_E_ARG;
  uint64_t S = 0, T = 0;
  for (uint32_t c = 0; c < ncols; ++c) S += bw[c];
  for (uint32_t r = 0; r < nsplit; ++r) {
    if (split[r] == 0 || split[r] > 64) return FMCS_E_ARG;
    T += split[r];
  }
  if (S != T) return FMCS_E_ARG;
  Split s(split, split + nsplit);
  *est_ms = estimate(m, (double)nrows, ncols, bw, card, col_order, s, truncate_cardinality != 0, parts4);
  return FMCS_OK;
}

namespace {
bool split_ok(uint32_t ncols, const uint32_t* bw, uint32_t nsplit, const uint32_t* split) {
  if (!split || nsplit == 0) return false;
  uint64_t S = 0, T = 0;
  for (uint32_t c = 0; c < ncols; ++c) S += bw[c];
  for (uint32_t r = 0; r < nsplit; ++r) {
    if (split[r] == 0 || split[r] > 64) return false;
    T += split[r];
  }
  return S == T;
}
bool order_ok(uint32_t ncols, const uint32_t* order) {
  if (!order) return false;
  std::vector<char> seen(ncols, 0);
  for (uint32_t i = 0; i < ncols; ++i) {
    if (order[i] >= ncols || seen[order[i]]) return false;
    seen[order[i]] = 1;
  }
  return true;
}
bool prims_ok(const fmcs_cost_primitives* p) { return p && p->sort_cost && p->stitch_cost && p->reconstruct_cost; }
}  // namespace

int fmcs_plan_statistics(uint32_t ncols, const uint32_t* bw, const uint64_t* card, const uint32_t* col_order,
                         uint32_t nsplit, const uint32_t* split, double* card_out, uint32_t* assemble9,
                         uint32_t* shift3) {
  if (!inputs_ok(ncols, bw, card) || !order_ok(ncols, col_order) || !split_ok(ncols, bw, nsplit, split))
    return FMCS_E_ARG;
  std::vector<double> cd;
  Stats st;
  plan_stats(ncols, bw, card, col_order, Split(split, split + nsplit), cd, st);
  if (card_out) std::copy(cd.begin(), cd.end(), card_out);
  if (assemble9) std::memcpy(assemble9, st.assemble, sizeof(st.assemble));
  if (shift3) std::memcpy(shift3, st.shift, sizeof(st.shift));
  return FMCS_OK;
}

int fmcs_plan_statistics_u64(uint32_t ncols, const uint32_t* bw, const uint64_t* card, const uint32_t* col_order,
                             uint32_t nsplit, const uint32_t* split, uint64_t* card_out, uint64_t* groups_out) {
  if (!inputs_ok(ncols, bw, card) || !order_ok(ncols, col_order) || !split_ok(ncols, bw, nsplit, split))
    return FMCS_E_ARG;
  std::vector<uint64_t> cu;
  Stats st;
  plan_stats(ncols, bw, card, col_order, Split(split, split + nsplit), cu, st);
  if (card_out) std::copy(cu.begin(), cu.end(), card_out);
  if (groups_out && !cu.empty()) {
    uint64_t acc = cu[0];
    groups_out[0] = 1;
    for (size_t r = 1; r < cu.size(); ++r) {
      groups_out[r] = acc;
      acc *= cu[r];
    }
  }
  return FMCS_OK;
}

int fmcs_estimate_group_num(uint32_t bit_position, uint32_t ncols, const uint32_t* bw, const uint64_t* card,
                            const uint32_t* col_order, double* groups) {
  if (!groups || !inputs_ok(ncols, bw, card) || !order_ok(ncols, col_order)) return FMCS_E_ARG;
  *groups = estimate_group_num(bit_position, ncols, bw, card, col_order);
  return FMCS_OK;
}

int fmcs_column_cost_prim(const fmcs_cost_primitives* p, uint64_t nrows, double group_rows, uint32_t bank_type,
                          double* cost) {
  if (!prims_ok(p) || !cost || bank_type > 2) return FMCS_E_ARG;
  *cost = single_column_cost_prim(p, (double)nrows, group_rows, (int)bank_type);
  return FMCS_OK;
}

int fmcs_estimate_plan_prim(const fmcs_cost_primitives* p, uint64_t nrows, uint32_t ncols, const uint32_t* bw,
                            const uint64_t* card, const uint32_t* col_order, uint32_t nsplit, const uint32_t* split,
                            int truncate, double* cost) {
  if (!prims_ok(p) || !cost || !inputs_ok(ncols, bw, card) || !order_ok(ncols, col_order) ||
      !split_ok(ncols, bw, nsplit, split))
    return FMCS_E_ARG;
  *cost = estimate_prim(p, (double)nrows, ncols, bw, card, col_order, Split(split, split + nsplit), truncate != 0);
  return FMCS_OK;
}

int fmcs_greedy_plans_prim(const fmcs_cost_primitives* p, uint64_t nrows, uint32_t ncols, const uint32_t* bw,
                           const uint64_t* card, const uint32_t* col_order, uint32_t rounds, uint32_t* out,
                           uint64_t cap_words, uint64_t* nplans, uint64_t* words) {
  if (!prims_ok(p) || rounds < 1 || rounds > 20 || !inputs_ok(ncols, bw, card) || !order_ok(ncols, col_order))
    return FMCS_E_ARG;
  auto cc = [p](double nn, double g, int bt) { return single_column_cost_prim(p, nn, g, bt); };
  return write_plans(greedy_plans(cc, (double)nrows, ncols, bw, card, col_order, rounds), out, cap_words, nplans,
                     words);
}

int fmcs_plan_enum_rrs_prim(const fmcs_cost_primitives* p, uint64_t nrows, uint32_t ncols, const uint32_t* bw,
                            const uint64_t* card, const uint32_t* col_order, uint32_t max_round, uint64_t seed,
                            uint32_t* out, uint64_t cap_words, uint64_t* nplans, uint64_t* words) {
  if (!prims_ok(p) || max_round == 0 || !inputs_ok(ncols, bw, card) || !order_ok(ncols, col_order))
    return FMCS_E_ARG;
  uint32_t S = 0;
  for (uint32_t c = 0; c < ncols; ++c) S += bw[c];
  if (max_round * 64u < S) return FMCS_E_ARG;
  const double n = (double)nrows;
  PlanEnumRRS e([=](const Split& sp) { return estimate_prim(p, n, ncols, bw, card, col_order, sp, true); }, S,
                max_round, seed);
  e.Optimize();
  std::vector<Split> v(e.Plans().begin(), e.Plans().end());
  return write_plans(v, out, cap_words, nplans, words);
}

int fmcs_split_exhaustive(uint32_t S, uint32_t rounds, uint32_t maxv, uint32_t* out, uint64_t cap_plans,
                          uint64_t* count) {
  if (!count) return FMCS_E_ARG;
  std::vector<Split> v = split_exhaustive(S, rounds, maxv);
  *count = v.size();
  if (out)
    for (uint64_t i = 0; i < v.size() && i < cap_plans; ++i)
      for (uint32_t r = 0; r < rounds; ++r) out[i * rounds + r] = v[i][r];
  return FMCS_OK;
}

int fmcs_enum_banksize_category(uint32_t rounds, uint32_t* out, uint64_t cap_plans, uint64_t* count) {
  if (!count || rounds == 0 || rounds > 20) return FMCS_E_ARG;
  std::vector<Split> v = bank_categories(rounds);
  *count = v.size();
  if (out)
    for (uint64_t i = 0; i < v.size() && i < cap_plans; ++i)
      for (uint32_t r = 0; r < rounds; ++r) out[i * rounds + r] = v[i][r];
  return FMCS_OK;
}

int fmcs_greedy_plans(const fmcs_cost_model* m, uint64_t nrows, uint32_t ncols, const uint32_t* bw,
                      const uint64_t* card, const uint32_t* col_order, uint32_t rounds, uint32_t* out,
                      uint64_t cap_words, uint64_t* nplans, uint64_t* words) {
  if (!m || !col_order || rounds < 1 || rounds > 20 || !inputs_ok(ncols, bw, card)) return FMCS_E_ARG;
  return write_plans(greedy_plans(m, (double)nrows, ncols, bw, card, col_order, rounds), out, cap_words, nplans,
                     words);
}

int fmcs_plan_enum_rrs(const fmcs_cost_model* m, uint64_t nrows, uint32_t ncols, const uint32_t* bw,
                       const uint64_t* card, const uint32_t* col_order, uint32_t max_round, uint64_t seed,
                       uint32_t* out, uint64_t cap_words, uint64_t* nplans, uint64_t* words) {
  if (!m || !col_order || max_round == 0 || !inputs_ok(ncols, bw, card)) return FMCS_E_ARG;
  uint32_t S = 0;
  for (uint32_t c = 0; c < ncols; ++c) S += bw[c];
  if (max_round * 64u < S) return FMCS_E_ARG;
  const double n = (double)nrows;
  PlanEnumRRS e([=](const Split& sp) { return estimate(m, n, ncols, bw, card, col_order, sp, true, nullptr); }, S,
                max_round, seed);
  e.Optimize();
  std::vector<Split> v(e.Plans().begin(), e.Plans().end());
  return write_plans(v, out, cap_words, nplans, words);
}

int fmcs_optimize(const fmcs_cost_model* m, uint64_t nrows, uint32_t ncols, const uint32_t* bw,
                  const uint64_t* card, int ordered, uint32_t* best_order, uint32_t* best_split,
                  uint32_t* best_nsplit, double* best_cost, uint64_t* nplans_costed) {
  if (!m || !best_order || !best_split || !best_nsplit || !best_cost || !inputs_ok(ncols, bw, card))
    return FMCS_E_ARG;
  uint32_t S = 0;
  for (uint32_t c = 0; c < ncols; ++c) S += bw[c];
  // column orders: n! permutations for GROUP BY (ordered == 0), in the
  // swap order of StitchComposer::permutations (src/stitch_composer.cpp:15-35)
  std::vector<Split> orders;
  Split base(ncols);
  for (uint32_t c = 0; c < ncols; ++c) base[c] = c;
  if (ordered || ncols > 8) {
    orders.push_back(base);
  } else {
    std::vector<uint32_t> arr = base;
    struct Rec {
      static void go(std::vector<uint32_t>& a, uint32_t k, uint32_t mm, std::vector<Split>& out) {
        if (k == mm) {
          out.push_back(a);
          return;
        }
        for (uint32_t i = k; i <= mm; ++i) {
          std::swap(a[k], a[i]);
          go(a, k + 1, mm, out);
          std::swap(a[k], a[i]);
        }
      }
    };
    Rec::go(arr, 0, ncols - 1, orders);
  }
  // max_round_num = ceil((2S - 2) / MIN_BANK_SIZE) (src/stitch_composer.cpp:3889-3890)
  uint32_t max_round = (uint32_t)std::ceil((2.0 * S - 2.0) / 16.0);
  max_round = std::max<uint32_t>(max_round, 1);
  double best = std::numeric_limits<double>::max();
  uint64_t costed = 0;
  auto t0 = std::chrono::steady_clock::now();
  double checkpoint = 0;
  for (const Split& ord : orders) {
    for (uint32_t r = 1; r <= max_round; ++r) {
      std::vector<Split> plans = r <= 2 ? split_exhaustive(S, r, 64)
                                        : greedy_plans(m, (double)nrows, ncols, bw, card, ord.data(), r);
      if (plans.empty()) continue;
      double loc = std::numeric_limits<double>::max();
      size_t li = 0;
      for (size_t i = 0; i < plans.size(); ++i) {
        double c = estimate(m, (double)nrows, ncols, bw, card, ord.data(), plans[i], false, nullptr);
        ++costed;
        if (c < loc) {
          loc = c;
          li = i;
        }
      }
      if (loc < best) {
        best = loc;
        std::copy(ord.begin(), ord.end(), best_order);
        std::copy(plans[li].begin(), plans[li].end(), best_split);
        *best_nsplit = (uint32_t)plans[li].size();
      }
      // stop once optimisation time exceeds THRESHOLD_PERC of the best
      // plan's cost (src/stitch_composer.cpp:3946-3953)
      double elapsed = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (elapsed - checkpoint > 0.01 * best) {
        checkpoint = elapsed;
        break;
      }
    }
  }
  *best_cost = best;
  if (nplans_costed) *nplans_costed = costed;
  return best < std::numeric_limits<double>::max() ? FMCS_OK : FMCS_E_ARG;
}

}  // extern "C"
