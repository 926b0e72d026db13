/*
 * fmcs_oracle.c -- CPU restatement of FastMultiColSort's multi-column sorting
 * path, used ONLY as the checker for the HIP implementation.
 *
 * TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library.  The product path never
 * calls it and has no CPU fallback.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Plain C99, scalar, single-threaded.  All arithmetic is integer.
 *
 * Pinning (what checks each function independently of this file):
 *   orc_construct_entries  - known answers from tests/misc_test.cpp:741-773
 *                            and :1263-1295, plus brute-force bit accounting
 *   orc_stitch             - bit_borrow known answer (0x123456, 0x78912345),
 *                            and the bijection property checked with numpy
 *   orc_sort_by_group /
 *   orc_multi_round_sort   - numpy lexsort of the original columns (massaging
 *                            is a bijection of the concatenated key), and the
 *                            reference itself when oracle/_ref builds
 *   orc_extract_group_info - numpy diff/cumsum closed form
 *   orc_subgroup_range /
 *   orc_multi_round_hash   - subgroup_test known answer [0,1,1,2,3,0,4,5,5]
 *                            (tests/subgroup_test.cpp:92-155) and a python
 *                            dict-based first-appearance restatement
 *   orc_split_exhaustive   - C(S-1, r-1) counts, descending-first order
 *   orc_hashmask           - closed form checks (power of two, 1/3 rule)
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint32_t ngroups;
    uint32_t sum_group_size;
    uint32_t ngroups_incache;
    uint32_t ngroups_outofcache;
    uint32_t ngroups_gtOne;
} orc_group_info;

/* ---- bank helpers ---------------------------------------------------- */

/* Bank width choice: src/stitch_composer.cpp:622-650 (sort path, never u8)
 * and :883-890 (hash path, allows u8). */
uint32_t orc_bank_bytes(uint32_t width, int allow_u8)
{
    if (allow_u8 && width <= 8) return 1;
    if (width <= 16) return 2;
    if (width <= 32) return 4;
    return 8;
}

static inline uint64_t ld(const void *p, uint32_t bank, uint64_t i)
{
    switch (bank) {
    case 1: return ((const uint8_t *)p)[i];
    case 2: return ((const uint16_t *)p)[i];
    case 4: return ((const uint32_t *)p)[i];
    default: return ((const uint64_t *)p)[i];
    }
}

static inline void st(void *p, uint32_t bank, uint64_t i, uint64_t v)
{
    switch (bank) {
    case 1: ((uint8_t *)p)[i] = (uint8_t)v; break;
    case 2: ((uint16_t *)p)[i] = (uint16_t)v; break;
    case 4: ((uint32_t *)p)[i] = (uint32_t)v; break;
    default: ((uint64_t *)p)[i] = v; break;
    }
}

static inline uint64_t lowmask(uint32_t w) { return w >= 64 ? ~0ULL : ((1ULL << w) - 1); }

/* ---- a1: construct-entry mapper -------------------------------------- */

/* Restates the Step-a loop of StitchComposer::run_single_plan
 * (src/stitch_composer.cpp:571-612): old columns (in col_order) are consumed
 * MSB-first into the new columns of `split`.  Ranges are 1-based from the
 * source MSB, inclusive.  Returns the number of entries, or -1 if the split
 * does not cover exactly sum(bw) bits or max_entries is too small. */
int orc_construct_entries(uint32_t ncols, const uint32_t *bw, const uint32_t *col_order,
                          uint32_t nsplit, const uint32_t *split,
                          uint32_t *ent_newcol, uint32_t *ent_src,
                          uint32_t *ent_start, uint32_t *ent_end, uint32_t max_entries)
{
    uint64_t sum_old = 0, sum_new = 0;
    for (uint32_t i = 0; i < ncols; ++i) sum_old += bw[col_order[i]];
    for (uint32_t i = 0; i < nsplit; ++i) { if (split[i] == 0) return -1; sum_new += split[i]; }
    if (sum_old != sum_new || ncols == 0 || nsplit == 0) return -1;
    uint32_t old_idx = 0, new_idx = 0, start = 1, n = 0;
    uint32_t old_w = bw[col_order[0]], new_w = split[0];
    while (new_idx < nsplit) {
        uint32_t m = old_w < new_w ? old_w : new_w;
        old_w -= m;
        new_w -= m;
        if (n >= max_entries) return -1;
        ent_newcol[n] = new_idx;
        ent_src[n] = col_order[old_idx];
        ent_start[n] = start;
        ent_end[n] = start + m - 1;
        ++n;
        start += m;
        if (old_w == 0) {
            ++old_idx;
            if (old_idx < ncols) old_w = bw[col_order[old_idx]];
            start = 1;
        }
        if (new_w == 0) {
            ++new_idx;
            if (new_idx < nsplit) new_w = split[new_idx];
        }
    }
    return (int)n;
}

/* ---- a3: stitch / bit-borrow ----------------------------------------- */

/* Restates Step b2 of run_single_plan (src/stitch_composer.cpp:652-703):
 * for every entry of a new column, ShiftLeftBatch(width) unless it is the
 * first entry (src/column.h:105-180), then AssembleFromColumn, i.e.
 * dst |= (src >> (bw_src - end)) & mask(end-start+1) (src/column.h:208-1059).
 * Entries must be grouped by ent_newcol in ascending order (as the mapper
 * emits them).  dst banks are written from zero (the reference memsets them,
 * src/stitch_composer.cpp:631-649). */
void orc_stitch(uint64_t n, const void *const *src, const uint32_t *src_bank, const uint32_t *src_bw,
                uint32_t nent, const uint32_t *ent_newcol, const uint32_t *ent_src,
                const uint32_t *ent_start, const uint32_t *ent_end,
                uint32_t ndst, void *const *dst, const uint32_t *dst_bank)
{
    for (uint32_t c = 0; c < ndst; ++c) {
        uint64_t dmask = dst_bank[c] >= 8 ? ~0ULL : ((1ULL << (8 * dst_bank[c])) - 1);
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t acc = 0;
            int first = 1;
            for (uint32_t e = 0; e < nent; ++e) {
                if (ent_newcol[e] != c) continue;
                uint32_t w = ent_end[e] - ent_start[e] + 1;
                uint32_t s = ent_src[e];
                if (!first) acc = (w >= 64) ? 0 : ((acc << w) & dmask);
                first = 0;
                uint64_t v = ld(src[s], src_bank[s], i);
                acc |= ((v >> (src_bw[s] - ent_end[e])) & lowmask(w)) & dmask;
            }
            st(dst[c], dst_bank[c], i, acc);
        }
    }
}

/* ---- a6: BATproject ---------------------------------------------------- */

/* Composer::BATproject (src/composer.cpp:608-628): out[i] = values[order[i]]. */
void orc_gather(uint32_t bank, uint64_t n, const void *values, const uint32_t *order, void *out)
{
    for (uint64_t i = 0; i < n; ++i) st(out, bank, i, ld(values, bank, order[i]));
}

/* ---- a9: extractGroupInfo -------------------------------------------- */

/* Composer::extractGroupInfo (src/composer.cpp:630-678): dense, monotone
 * group ids; a new id starts where the value or the previous group changes.
 * Returns the number of groups. */
uint32_t orc_extract_group_info(uint32_t bank, uint64_t n, const void *sorted_values,
                                const uint32_t *in_group, uint32_t *out_group)
{
    if (n == 0) return 0;
    uint32_t cur = 0;
    out_group[0] = 0;
    uint64_t prev = ld(sorted_values, bank, 0);
    for (uint64_t i = 1; i < n; ++i) {
        uint64_t v = ld(sorted_values, bank, i);
        cur += (v != prev) || (in_group && in_group[i] != in_group[i - 1]);
        prev = v;
        out_group[i] = cur;
    }
    return cur + 1;
}

/* ---- a7/a8: per-segment stable sort ----------------------------------- */

/* The reference sorts each segment with an unstable AVX2 merge sort
 * (src/mergesort.h:2878-3096) whose tie order is not part of the contract.
 * The oracle fixes ties by oid: a stable merge sort by key, which equals a
 * sort by (key, oid) because within a segment the incoming rows are in
 * ascending oid order (round 1 starts from the identity order,
 * src/composer.cpp:60-63, and every later segment holds rows that tie on all
 * earlier keys). */
static void msort(uint64_t *k, uint32_t *o, uint64_t *tk, uint32_t *to, uint64_t n)
{
    if (n < 2) return;
    if (n <= 16) {
        for (uint64_t i = 1; i < n; ++i) {
            uint64_t kk = k[i];
            uint32_t oo = o[i];
            uint64_t j = i;
            while (j > 0 && k[j - 1] > kk) { k[j] = k[j - 1]; o[j] = o[j - 1]; --j; }
            k[j] = kk;
            o[j] = oo;
        }
        return;
    }
    uint64_t h = n / 2;
    msort(k, o, tk, to, h);
    msort(k + h, o + h, tk + h, to + h, n - h);
    uint64_t a = 0, b = h, t = 0;
    while (a < h && b < n) {
        if (k[b] < k[a]) { tk[t] = k[b]; to[t++] = o[b++]; }
        else { tk[t] = k[a]; to[t++] = o[a++]; }
    }
    while (a < h) { tk[t] = k[a]; to[t++] = o[a++]; }
    while (b < n) { tk[t] = k[b]; to[t++] = o[b++]; }
    memcpy(k, tk, n * sizeof(uint64_t));
    memcpy(o, to, n * sizeof(uint32_t));
}

/* Composer::subsort (src/composer.cpp:558-589): segments of <= 1 row are
 * left alone. */
static void subsort(uint64_t *k, uint32_t *o, uint64_t *tk, uint32_t *to, uint64_t n)
{
    if (n <= 1) return;
    msort(k, o, tk, to, n);
}

static uint32_t blocksize_of(uint32_t bank)
{
    /* BLOCKSIZE<T>() = L2 / (2*2*max(sizeof(oid), sizeof(T))), L2 = 256 KiB
     * (src/params.h:141-143, :1238 region). */
    uint32_t w = bank > 4 ? bank : 4;
    return (256u * 1024u) / (4u * w);
}

/* Composer::sortByGroup<T> (src/composer.cpp:1222-1408).
 *  column    - the round's bank, indexed by oid (the reference's inValues)
 *  order     - in: order after the previous round (identity for round 1);
 *              out: order after this round
 *  in_group  - NULL for round 1, else the previous round's group ids
 *  out_group - written unless last_round (extractGroupInfo is skipped)
 *  sorted_out- optional: the sorted values of this round
 * Returns the number of output groups (0 on the last round).  info receives
 * the group statistics the reference accumulates for rounds > 1
 * (src/composer.cpp:1308-1332). */
uint32_t orc_sort_by_group(uint32_t bank, uint64_t n, const void *column, uint32_t *order,
                           const uint32_t *in_group, uint32_t *out_group, int last_round,
                           void *sorted_out, orc_group_info *info)
{
    if (n == 0) return 0;
    uint64_t *k = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *tk = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint32_t *to = (uint32_t *)malloc(n * sizeof(uint32_t));
    /* round 1 sorts the column in place (identity order); later rounds
     * gather it through `order` first (BATproject, :1257-1261). */
    for (uint64_t i = 0; i < n; ++i) k[i] = ld(column, bank, order[i]);
    uint32_t bs = blocksize_of(bank);
    if (in_group) {
        uint64_t start = 0;
        for (uint64_t end = 1; end <= n; ++end) {
            if (end == n || in_group[end] != in_group[start]) {
                uint64_t sz = end - start;
                subsort(k + start, order + start, tk, to, sz);
                if (info) {
                    info->ngroups++;
                    info->sum_group_size += (uint32_t)sz;
                    if (sz < bs) info->ngroups_incache++; else info->ngroups_outofcache++;
                    info->ngroups_gtOne += (sz > 1);
                }
                start = end;
            }
        }
    } else {
        subsort(k, order, tk, to, n);
    }
    uint32_t ng = 0;
    if (sorted_out)
        for (uint64_t i = 0; i < n; ++i) st(sorted_out, bank, i, k[i]);
    if (!last_round && out_group) {
        uint32_t cur = 0;
        out_group[0] = 0;
        for (uint64_t i = 1; i < n; ++i) {
            cur += (k[i] != k[i - 1]) || (in_group && in_group[i] != in_group[i - 1]);
            out_group[i] = cur;
        }
        ng = cur + 1;
    }
    free(k);
    free(tk);
    free(to);
    return ng;
}

/* Composer::multiRoundSorting (src/composer.cpp:48-135), but the results are
 * kept: order (final oid permutation), and for every round r < nrounds-1 the
 * group ids in groups[r*n ...] (may be NULL).  Returns the last computed
 * group count. */
uint32_t orc_multi_round_sort(uint32_t nrounds, const void *const *cols, const uint32_t *banks,
                              uint64_t n, uint32_t *order, uint32_t *groups,
                              orc_group_info *info)
{
    for (uint64_t i = 0; i < n; ++i) order[i] = (uint32_t)i;
    uint32_t *gin = NULL;
    uint32_t *ga = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *gb = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t ng = 0;
    for (uint32_t r = 0; r < nrounds; ++r) {
        int last = (r == nrounds - 1);
        uint32_t *gout = (gin == ga) ? gb : ga;
        uint32_t g = orc_sort_by_group(banks[r], n, cols[r], order, gin, gout, last, NULL, info);
        if (!last) {
            ng = g;
            if (groups) memcpy(groups + (uint64_t)r * n, gout, n * sizeof(uint32_t));
            gin = gout;
        }
    }
    free(ga);
    free(gb);
    return ng;
}

/* ---- a10/a11/a12: hash grouping --------------------------------------- */


/* Hashing::HASHmask (src/hashing.h:442-465). */
uint64_t orc_hashmask(uint64_t cnt)
{
    uint64_t m = cnt;
    m |= m >> 1; m |= m >> 2; m |= m >> 4; m |= m >> 8; m |= m >> 16; m |= m >> 32;
    m -= m >> 1;
    if (m + m - cnt < 2 * (cnt - m)) m += m;
    return m;
}

/* Hashing::subgroup_range (src/hashing.h:289-385): dense ids in order of
 * first appearance within rows [start, start+nitems) of `values`, numbered
 * from start_gid.  The chained table is replaced by a sorted index, which
 * gives the same ids (the ids depend only on first appearance).  This is the
 * intended semantics; the reference's own block loop reuses rem_sz for every
 * block (:345) and so is only faithful for nitems <= one block.
 * Returns the number of groups in the range. */
uint32_t orc_subgroup_range(uint32_t bank, const void *values, uint64_t start, uint64_t nitems,
                            uint32_t start_gid, uint32_t *gid_out)
{
    if (nitems == 0) return 0;
    /* sort (value, first index) pairs to find each row's first occurrence */
    uint64_t *k = (uint64_t *)malloc(nitems * sizeof(uint64_t));
    uint32_t *o = (uint32_t *)malloc(nitems * sizeof(uint32_t));
    uint64_t *tk = (uint64_t *)malloc(nitems * sizeof(uint64_t));
    uint32_t *to = (uint32_t *)malloc(nitems * sizeof(uint32_t));
    uint32_t *first = (uint32_t *)malloc(nitems * sizeof(uint32_t));
    for (uint64_t i = 0; i < nitems; ++i) { k[i] = ld(values, bank, start + i); o[i] = (uint32_t)i; }
    msort(k, o, tk, to, nitems); /* stable: equal values keep ascending index */
    uint64_t j = 0;
    while (j < nitems) {
        uint64_t e = j;
        while (e < nitems && k[e] == k[j]) ++e;
        for (uint64_t t = j; t < e; ++t) first[o[t]] = o[j];
        j = e;
    }
    /* ids by first appearance */
    uint32_t *idof = to; /* reuse: id of a first-occurrence row */
    uint32_t g = start_gid;
    for (uint64_t i = 0; i < nitems; ++i) {
        if (first[i] == i) idof[i] = g++;
        gid_out[start + i] = idof[first[i]];
    }
    free(k); free(o); free(tk); free(to); free(first);
    return g - start_gid;
}

/* Partition::histogram_based_reordering (src/partition.h:44-161): stable
 * counting sort of the oids in [start, start+nitems) by their group id
 * (ids in [start_gid, start_gid+ngroups)); writes the clustered outGroup and
 * outOrder for that range. */
void orc_histogram_reorder(uint64_t start, uint64_t nitems, const uint32_t *in_oid,
                           const uint32_t *gid, uint32_t start_gid, uint32_t ngroups,
                           uint32_t *out_group, uint32_t *out_order)
{
    uint32_t *hist = (uint32_t *)calloc(ngroups ? ngroups : 1, sizeof(uint32_t));
    for (uint64_t i = 0; i < nitems; ++i) hist[gid[start + i] - start_gid]++;
    uint64_t acc = start;
    for (uint32_t g = 0; g < ngroups; ++g) {
        uint32_t c = hist[g];
        for (uint32_t t = 0; t < c; ++t) out_group[acc + t] = start_gid + g;
        hist[g] = (uint32_t)(acc - start);
        acc += c;
    }
    for (uint64_t i = 0; i < nitems; ++i) {
        uint32_t g = gid[start + i] - start_gid;
        out_order[start + hist[g]++] = in_oid[start + i];
    }
    free(hist);
}

/* Composer::multiRoundHashing_reordered with hashByGroup_reordered
 * (src/composer.cpp:137-227, :814-976): per round, gather the bank through
 * the current order (not in round 1), give first-appearance ids per
 * segment of the previous round's groups, then (unless it is the last round)
 * cluster the oids by id.  Outputs:
 *   final_gid  - the last round's ids, per row position (the reference's
 *                intermediate_group_idx after the last round)
 *   order      - the order the last round hashed in
 *   round_groups (optional) - clustered group ids after each non-last round
 * Returns the number of groups after the last round. */
uint32_t orc_multi_round_hash(uint32_t nrounds, const void *const *cols, const uint32_t *banks,
                              uint64_t n, uint32_t *order, uint32_t *final_gid,
                              uint32_t *round_groups)
{
    uint32_t *in_order = order;
    uint32_t *out_order = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *gid = final_gid;
    uint32_t *in_group = NULL;
    uint32_t *ga = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *gb = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint64_t *vals = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) in_order[i] = (uint32_t)i;
    uint32_t total = 0;
    for (uint32_t r = 0; r < nrounds; ++r) {
        int last = (r == nrounds - 1);
        for (uint64_t i = 0; i < n; ++i) vals[i] = ld(cols[r], banks[r], in_group ? in_order[i] : i);
        uint32_t *out_group = (in_group == ga) ? gb : ga;
        total = 0;
        uint64_t start = 0;
        for (uint64_t end = 1; end <= n; ++end) {
            if (end == n || (in_group && in_group[end] != in_group[start])) {
                if (!in_group && end != n) continue;
                uint32_t g = orc_subgroup_range(8, vals, start, end - start, total, gid);
                if (!last) orc_histogram_reorder(start, end - start, in_order, gid, total, g, out_group, out_order);
                total += g;
                start = end;
            }
        }
        if (!last) {
            if (round_groups) memcpy(round_groups + (uint64_t)r * n, out_group, n * sizeof(uint32_t));
            memcpy(in_order, out_order, n * sizeof(uint32_t));
            in_group = out_group;
        }
    }
    free(out_order); free(ga); free(gb); free(vals);
    return total;
}

/* ---- a13: split enumeration ------------------------------------------ */

static void split_rec(uint32_t n, uint32_t cur, uint32_t rounds, uint32_t maxv, uint32_t *part,
                      uint32_t *out, uint64_t *count, uint64_t cap)
{
    if (cur == rounds - 1) {
        if (n > maxv) return;
        part[cur] = n;
        if (*count < cap)
            for (uint32_t i = 0; i < rounds; ++i) out[*count * rounds + i] = part[i];
        ++*count;
        return;
    }
    for (uint32_t i = n > maxv ? maxv : n; i > 0; --i) {
        part[cur] = i;
        if (n - i > 0) split_rec(n - i, cur + 1, rounds, maxv, part, out, count, cap);
    }
}

/* StitchComposer::split_exhaustive_certain_round (src/stitch_composer.cpp:
 * 205-230): every composition of S into `rounds` parts each <= maxv, first
 * part descending.  Writes up to cap plans (rounds values each) and returns
 * the total count. */
uint64_t orc_split_exhaustive(uint32_t S, uint32_t rounds, uint32_t maxv, uint32_t *out, uint64_t cap)
{
    uint32_t part[64];
    uint64_t count = 0;
    if (rounds == 0 || rounds > 64 || S == 0) return 0;
    split_rec(S, 0, rounds, maxv, part, out, &count, cap);
    return count;
}
