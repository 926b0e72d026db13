/*
 * fmcs_plan.h -- C ABI of the host-side massaging planner and of the
 * composer-level executor (run one plan end to end on the GPU).
 *
 * The planner keeps the reference's structure (src/stitch_composer.cpp,
 * src/plan_enum_rrs.cpp, src/rrs.h); its cost constants are per-row GPU
 * times (fmcs_cost_model) instead of the reference's CPU cycle tables
 * (src/params.h:200-2283).  Plans are splits: new round widths, MSB first,
 * summing to the total key width.  Plan lists are returned flattened as
 * [len, w0, w1, ...] records.
 */
#ifndef FMCS_PLAN_H
#define FMCS_PLAN_H

#include <stdint.h>

#include "fmcs.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FMCS_COST_GBUCKETS 32

/* Per-row GPU costs in nanoseconds, by bank type (0: u16, 1: u32, 2: u64). */
typedef struct {
  double radix_pass_ns[3];                     /* one LSD pass of the round-1 sort */
  double segsort_ns[3][FMCS_COST_GBUCKETS];    /* later rounds, by log2(avg group size) */
  double gather_ns[3];                         /* oid gather (tuple reconstruction) */
  double boundary_ns[3];                       /* group boundary extraction */
  double stitch_ns_per_byte;                   /* fused stitch, per byte moved */
  double round_overhead_ms;                    /* fixed cost per round */
  double pass_overhead_ms;                     /* fixed cost per round-1 LSD pass */
  double level_overhead_ms;                    /* fixed cost per level of a later round's segmented sort
                                                  (one level per 8 key bits, plus one) */
  uint32_t carry_next;                         /* fmcs_tuning.carry_next of the executor: 1 and 2 price
                                                  round 2's column as carried through round 1's radix
                                                  passes (1: packed into u64 composites where it fits),
                                                  no gather before round 2; 0: as gathered */
} fmcs_cost_model;

int fmcs_cost_model_default(fmcs_cost_model* m);

/* Cost primitives supplied by the caller, for the planner's structure with
 * someone else's constants -- for instance the reference's own estimators
 * (StitchComposer::estimate_mergesort_uint{16,32,64}, estimate_stitch,
 * estimate_tuple_reconstruct; src/stitch_composer.h:54-72), so that a CPU
 * executor keeps its cost tables and this planner prices plans exactly as
 * the reference does.  Units are the caller's. */
typedef struct {
  /* cost of sorting one group of group_rows rows in a bank of type bank_type (0: u16, 1: u32, 2: u64) */
  double (*sort_cost)(void* user, double group_rows, uint32_t bank_type);
  /* stitch cost from assemble[from bank][to bank] (row-major, 9 words) and shift[to bank] entry counts */
  double (*stitch_cost)(void* user, const uint32_t* assemble9, const uint32_t* shift3, uint64_t nrows);
  /* tuple-reconstruction cost of a plan's new banks (bytes per round, rounds 1.. gathered) */
  double (*reconstruct_cost)(void* user, const uint32_t* bytewidths, uint32_t nrounds, uint64_t nrows);
  void* user;
} fmcs_cost_primitives;

/* getCardinalityFromStatistics (src/stitch_composer.cpp:288-553): the
 * cardinality of every new column but the last (card_per_bit^width products,
 * card_per_bit = card^(1/bw)), card_out[nsplit - 1] = 1, and the assemble /
 * shift statistics of the split's construct entries. */
int fmcs_plan_statistics(uint32_t ncols, const uint32_t* bw, const uint64_t* card, const uint32_t* col_order,
                         uint32_t nsplit, const uint32_t* split, double* card_out, uint32_t* assemble9,
                         uint32_t* shift3);
/* PlanEnumRRS::GetStatistics (src/plan_enum_rrs.cpp:269-396): the same
 * cardinalities held in uint64 and converted after every entry (x86-64
 * conversion), and groups_out (optional) the groups before each round,
 * their uint64 products with wrap-around (src/plan_enum_rrs.cpp:149-156). */
int fmcs_plan_statistics_u64(uint32_t ncols, const uint32_t* bw, const uint64_t* card, const uint32_t* col_order,
                             uint32_t nsplit, const uint32_t* split, uint64_t* card_out, uint64_t* groups_out);
/* estimateGroupNum (src/stitch_composer.cpp:1845-1867): groups of the first
 * bit_position bits of the ordered columns. */
int fmcs_estimate_group_num(uint32_t bit_position, uint32_t ncols, const uint32_t* bw, const uint64_t* card,
                            const uint32_t* col_order, double* groups);
/* calculate_single_column_cost (src/stitch_composer.cpp:2186-2226): (n / g)
 * groups x f(g) x sort_cost(g), f(g) = 1 / (1 + 1/(g - 1)), 0 for g <= 1. */
int fmcs_column_cost_prim(const fmcs_cost_primitives* p, uint64_t nrows, double group_rows, uint32_t bank_type,
                          double* cost);
/* One plan's cost with the primitives: estimate_plan_collection_greedy's
 * costing (truncate 0; src/stitch_composer.cpp:2228-2378) or
 * PlanEnumRRS::GetCost's (truncate 1, uint64 cardinalities;
 * src/plan_enum_rrs.cpp:94-197). */
int fmcs_estimate_plan_prim(const fmcs_cost_primitives* p, uint64_t nrows, uint32_t ncols, const uint32_t* bw,
                            const uint64_t* card, const uint32_t* col_order, uint32_t nsplit, const uint32_t* split,
                            int truncate, double* cost);
/* fmcs_greedy_plans with the primitives' column cost. */
int fmcs_greedy_plans_prim(const fmcs_cost_primitives* p, uint64_t nrows, uint32_t ncols, const uint32_t* bw,
                           const uint64_t* card, const uint32_t* col_order, uint32_t rounds, uint32_t* out,
                           uint64_t cap_words, uint64_t* nplans, uint64_t* words);
/* fmcs_plan_enum_rrs costed by fmcs_estimate_plan_prim (truncate 1): with the
 * reference's estimators and the seed the reference's RRS took from
 * time(0) (src/rrs.h:68), the same visited plans as PlanEnumRRS. */
int fmcs_plan_enum_rrs_prim(const fmcs_cost_primitives* p, uint64_t nrows, uint32_t ncols, const uint32_t* bw,
                            const uint64_t* card, const uint32_t* col_order, uint32_t max_round, uint64_t seed,
                            uint32_t* out, uint64_t cap_words, uint64_t* nplans, uint64_t* words);

/* Estimated cost of one plan (estimate_plan_collection_greedy's per-plan
 * costing, src/stitch_composer.cpp:2228-2378; truncate_cardinality selects
 * PlanEnumRRS::GetCost's uint64 cardinalities, src/plan_enum_rrs.cpp:129-153).
 * parts4 (optional): stitch, tuple reconstruction, sort, group extraction. */
int fmcs_estimate_plan(const fmcs_cost_model* m, uint64_t nrows, uint32_t ncols, const uint32_t* bw,
                       const uint64_t* card, const uint32_t* col_order, uint32_t nsplit, const uint32_t* split,
                       int truncate_cardinality, double* est_ms, double* parts4);

/* split_exhaustive_certain_round (src/stitch_composer.cpp:205-230): all
 * compositions of S into `rounds` parts <= maxv, descending-first; writes up
 * to cap_plans plans of `rounds` values, *count = total. */
int fmcs_split_exhaustive(uint32_t S, uint32_t rounds, uint32_t maxv, uint32_t* out, uint64_t cap_plans,
                          uint64_t* count);

/* enum_banksize_category (src/stitch_composer.cpp:1717-1734): the 3^rounds
 * bank patterns (0: 16, 1: 32, 2: 64 bit), lexicographic. */
int fmcs_enum_banksize_category(uint32_t rounds, uint32_t* out, uint64_t cap_plans, uint64_t* count);

/* find_greedy_plans (src/stitch_composer.cpp:1992-2125): one plan per valid
 * bank pattern. */
int fmcs_greedy_plans(const fmcs_cost_model* m, uint64_t nrows, uint32_t ncols, const uint32_t* bw,
                      const uint64_t* card, const uint32_t* col_order, uint32_t rounds, uint32_t* out,
                      uint64_t cap_words, uint64_t* nplans, uint64_t* words);

/* PlanEnumRRS + RRS::Optimize (src/plan_enum_rrs.cpp:37-267, src/rrs.h:80-170)
 * with an explicit seed (the reference seeds with time(0), src/rrs.h:68):
 * the set of visited plans, ascending. */
int fmcs_plan_enum_rrs(const fmcs_cost_model* m, uint64_t nrows, uint32_t ncols, const uint32_t* bw,
                       const uint64_t* card, const uint32_t* col_order, uint32_t max_round, uint64_t seed,
                       uint32_t* out, uint64_t cap_words, uint64_t* nplans, uint64_t* words);

/* StitchComposer::optimizeAndRun's search (src/stitch_composer.cpp:
 * 3839-3985): column orders (all permutations when ordered == 0), rounds
 * 1..ceil((2S-2)/16), exhaustive for <= 2 rounds, greedy beyond, early stop
 * at 1% of the best cost.  Returns the best order and split. */
int fmcs_optimize(const fmcs_cost_model* m, uint64_t nrows, uint32_t ncols, const uint32_t* bw,
                  const uint64_t* card, int ordered, uint32_t* best_order, uint32_t* best_split,
                  uint32_t* best_nsplit, double* best_cost, uint64_t* nplans_costed);

/* ---- composer-level execution (needs a context) -------------------- */

/* StitchComposer::run_single_plan (src/stitch_composer.cpp:555-806):
 * construct entries, stitch the columns into new banks, then sort all
 * rounds (ChainComposer::SortAllColumns -> multiRoundSorting).  order_d
 * receives the final oid permutation; groups_d (optional) the final group
 * ids; ms6 (optional) the reference's breakdown: [0] stitching,
 * [1] tuple reconstruct, [2] order/group info, [3] all sorting rounds,
 * [4] first-round sort, [5] group extraction; *total_ms the whole plan. */
int fmcs_run_plan(fmcs_ctx* ctx, uint32_t ncols, const void* const* cols_d, const uint32_t* banks,
                  const uint32_t* bw, const uint32_t* col_order, uint32_t nsplit, const uint32_t* split, int mode,
                  uint64_t nrows, uint32_t* order_d, uint32_t* groups_d, double* ms6, double* total_ms,
                  uint32_t* ngroups);

/* fmcs_run_plan with a sort direction per column (descending[i] != 0:
 * descending, through fmcs_stitch_dir; NULL: all ascending).  Ties keep the
 * ascending oid order.  The reference sorts ascending only
 * (src/composer.cpp:561-563). */
int fmcs_run_plan_dir(fmcs_ctx* ctx, uint32_t ncols, const void* const* cols_d, const uint32_t* banks,
                      const uint32_t* bw, const uint8_t* descending, const uint32_t* col_order, uint32_t nsplit,
                      const uint32_t* split, int mode, uint64_t nrows, uint32_t* order_d, uint32_t* groups_d,
                      double* ms6, double* total_ms, uint32_t* ngroups);

/* StitchComposer::run_single_plan_hash (src/stitch_composer.cpp:808-1079):
 * the same with u8 banks allowed and hash grouping per round. */
int fmcs_run_plan_hash(fmcs_ctx* ctx, uint32_t ncols, const void* const* cols_d, const uint32_t* banks,
                       const uint32_t* bw, const uint32_t* col_order, uint32_t nsplit, const uint32_t* split,
                       int hash_type, uint64_t nrows, uint32_t* order_d, uint32_t* gid_d, double* total_ms,
                       uint32_t* ngroups);

/* StitchComposer::optimizeAndRun (src/stitch_composer.cpp:3839-4013): the
 * plan search of fmcs_optimize (column orders: all permutations when
 * ordered == 0, the GROUP BY case), costed by m (NULL: the default model;
 * fmcs_calibrate gives the device's), then run_single_plan of the best plan,
 * as fmcs_run_plan.  card: distinct values per column, or NULL to count them
 * on the device (fmcs_distinct_count).  best_order (ncols entries),
 * best_split (64 entries) and *best_nsplit receive the plan; *opt_ms the
 * host time of the search (and of the counting), ms6 / *total_ms the plan's
 * breakdown and GPU time. */
int fmcs_optimize_and_run(fmcs_ctx* ctx, const fmcs_cost_model* m, uint32_t ncols, const void* const* cols_d,
                          const uint32_t* banks, const uint32_t* bw, const uint64_t* card, int ordered, int mode,
                          uint64_t nrows, uint32_t* order_d, uint32_t* groups_d, uint32_t* best_order,
                          uint32_t* best_split, uint32_t* best_nsplit, double* opt_ms, double* ms6, double* total_ms,
                          uint32_t* ngroups);

/* Measures the cost model on this device with synthetic banks of nrows rows
 * (radix passes, gather, boundary, stitch, and segmented sorts over
 * groups of 2^k rows). */
int fmcs_calibrate(fmcs_ctx* ctx, uint64_t nrows, fmcs_cost_model* out);

/* ---- learned plan-cost model (no counterpart in the reference; its plan
 *      costs are analytic, src/stitch_composer.cpp:2186-2378) ---------- */

/* fp32 MLP nin -> h1 -> h2 -> 1 with ReLU; each of nin, h1, h2 <= 64.
 * Parameters (device, floats): W1[h1][nin], b1[h1], W2[h2][h1], b2[h2],
 * W3[h2], b3. */
typedef struct {
  uint32_t nin, h1, h2;
} fmcs_mlp_shape;

int fmcs_mlp_param_count(const fmcs_mlp_shape* s, uint32_t* count);

/* y_d[s] = MLP(x_d[s, 0..nin)); acts_d (optional, batch x (h1+h2)) keeps the
 * post-ReLU activations for the backward pass. */
int fmcs_mlp_forward(fmcs_ctx* ctx, const fmcs_mlp_shape* s, const float* params_d, const float* x_d,
                     uint32_t batch, float* y_d, float* acts_d);

/* Gradient of L = sum_s 0.5 (y_s - target_s)^2 / batch into grad_d (one
 * float per parameter), L into loss_d (optional).  Deterministic. */
int fmcs_mlp_backward(fmcs_ctx* ctx, const fmcs_mlp_shape* s, const float* params_d, const float* x_d,
                      const float* acts_d, const float* y_d, const float* target_d, uint32_t batch, float* grad_d,
                      float* loss_d);

/* One Adam step (bias-corrected, step >= 1) over nparams parameters. */
int fmcs_mlp_adam(fmcs_ctx* ctx, uint32_t nparams, float* params_d, const float* grad_d, float* m_d, float* v_d,
                  float lr, float beta1, float beta2, float eps, uint32_t step);

#ifdef __cplusplus
}
#endif

#endif /* FMCS_PLAN_H */
