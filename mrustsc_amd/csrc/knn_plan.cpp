// Device placement and launch order of the kNN pipeline for a set of embeddings.
//
// A plan lays every dataset's workspace out in one arena (coordinates, norms, column
// sums, packed MFMA operands, candidate buffers, kept candidates, thresholds) and the
// launch metadata in a second block (descriptors, persistent-pass work items, the list of
// rows for the exact kernel, scalars). Enqueueing issues the same launches for one
// embedding (with the pass kernel specialised at compile time) or for a batch (one
// persistent pass launch over every dataset's work items).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <type_traits>

#include "knn_launch.hpp"
#include "scrust_hip_internal.hpp"

namespace scrust_hip {

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

constexpr uint32_t kMinPiecePairs = 4;   // shortest reference range of a work item (tile pairs)
constexpr uint32_t kSelectKeysMax = 1024;  // keys per row the final selection takes (16 per lane)
constexpr uint32_t kTailSplits = 3;        // reference ranges of the blocks of a launch's last round

uint32_t next_pow2(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Reference ranges per query block for one launch of `blocks` blocks on `slots`
// workgroup slots: the count minimising rounds x (range length + fixed cost), where an
// item's fixed cost (query fragments, the first compactions) is worth `overhead` tile
// pairs, taking more ranges only for a predicted gain of 15% (the model leaves out the
// cache traffic of shorter ranges). Ranges are whole pairs, at least kMinPiecePairs long.
uint32_t choose_splits(uint32_t blocks, uint32_t pairs, uint64_t slots, uint32_t max_splits,
                       double overhead) {
    uint32_t best = 1;
    double one = 0.0, best_time = 0.0;
    for (uint32_t sp = 1; sp <= max_splits; sp++) {
        const uint32_t len = (pairs + sp - 1) / sp;
        if (sp > 1 && len < kMinPiecePairs) break;
        const uint32_t used = (pairs + len - 1) / len;
        if (used != sp) continue;
        const uint64_t rounds = (static_cast<uint64_t>(blocks) * sp + slots - 1) / slots;
        const double time = static_cast<double>(rounds) * (len + overhead);
        if (sp == 1) {
            one = best_time = time;
        } else if (time < one * 0.85 && time < best_time) {
            best = sp;
            best_time = time;
        }
    }
    return best;
}

// Every block's scan cut into `pieces` near-equal ranges (batched launches).
uint32_t partition_even(uint32_t set, uint32_t blocks, uint32_t pairs, uint32_t pieces,
                        std::vector<PassItem>& items) {
    pieces = std::max(1u, std::min(pieces, pairs));
    const uint32_t len = (pairs + pieces - 1) / pieces;
    pieces = (pairs + len - 1) / len;
    for (uint32_t b = 0; b < blocks; b++)
        for (uint32_t j = 0; j < pieces; j++)
            items.push_back(PassItem{set, b, j, 2 * j * len, 2 * std::min(pairs, (j + 1) * len), 0});
    return pieces;
}

}  // namespace

scrust_hip_status plan_layout(const std::vector<SetSpec>& specs,
                              const scrust_hip_knn_options* options, Plan& plan,
                              int compute_units, bool single, const TuneEntry* force,
                              const BatchKnobs* knobs, bool shared) {
    plan = Plan{};
    if (single && specs.size() != 1)
        return fail(SCRUST_HIP_KERNEL, "HIP kernel knn: a single call lays out exactly one dataset");
    const int metric = options ? static_cast<int>(options->metric) : kMetricEuclidean;
    const int precision = options ? static_cast<int>(options->precision) : SCRUST_HIP_BF16;
    uint32_t cap = options && options->tile ? options->tile : 256u;
    if (cap != 128 && cap != 256 && cap != 512)
        return invalid("tile", "0, 128, 256 or 512", std::to_string(cap));
    plan.precision = precision;
    plan.operand = precision == SCRUST_HIP_BF16 ? kOperandBf16
                   : precision == SCRUST_HIP_FP32 ? kOperandF32
                                                  : kOperandNone;
    if (precision == SCRUST_HIP_FP16) {
        // The local-origin fp16 pass needs the rows grouped (each 512-row work item's origin
        // is the mean of nearby rows): single datasets in spatial order, automatic from
        // kOrderMinCells cells or asked for. Anything else takes the bf16 split pass, which
        // gives the same graph.
        const uint32_t order_knob = force && force->cap ? force->order : (options ? options->order : 0u);
        const bool local = single && specs[0].k <= static_cast<size_t>(kMaxLocalK) &&
                           local_chunks(static_cast<int>(specs[0].d)) > 0 &&
                           (order_knob == 2u || (order_knob == 0u && specs[0].n >= kOrderMinCells));
        plan.operand = local ? kOperandF16Local : kOperandBf16;
        plan.precision = local ? SCRUST_HIP_FP16 : SCRUST_HIP_BF16;
    }
    // The packed 16-bit operands hold up to kMaxPackedD features: a wider single dataset
    // takes the fp32 pass (up to kMaxPassD features), which gives the same graph.
    if (single && plan.operand == kOperandBf16 && specs[0].d > static_cast<size_t>(kMaxPackedD) &&
        specs[0].d <= static_cast<size_t>(kMaxPassD)) {
        plan.operand = kOperandF32;
        plan.precision = SCRUST_HIP_FP32;
    }
    plan.shape.count = static_cast<uint32_t>(specs.size());
    plan.shape.metric = metric;
    plan.shape.operand = plan.operand;

    // Workgroup slots of the pass (two per CU) and, for batches, the longest reference
    // range of one item: by default about a third of a slot's share of the whole batch.
    const uint64_t slots = 2ull * static_cast<uint64_t>(compute_units > 0 ? compute_units : 256);
    uint64_t batch_pairs = 0;
    for (const SetSpec& sp : specs)
        batch_pairs += ((sp.n + kPassRows - 1) / kPassRows) * ((sp.n + 2 * kTileRefsHost - 1) / (2 * kTileRefsHost));
    const uint64_t cut = knobs && knobs->cut ? knobs->cut : 3u;
    const uint64_t batch_len = std::max<uint64_t>(kMinPiecePairs, (batch_pairs + cut * slots - 1) / (cut * slots));
    if (!single && knobs && knobs->cap) {
        if (knobs->cap != 128 && knobs->cap != 256 && knobs->cap != 512)
            return invalid("tile", "0, 128, 256 or 512", std::to_string(knobs->cap));
        cap = knobs->cap;
    }

    size_t arena = 0;
    size_t count_words = 0;  // segment counts of every dataset, one region (one clear)
    uint32_t row = 0, prep = 0;
    for (const SetSpec& sp : specs) {
        DatasetDesc s{};
        s.x = sp.x;
        s.out_indices = sp.out_indices;
        s.out_distances = sp.out_distances;
        s.n = static_cast<uint32_t>(sp.n);
        s.d = static_cast<uint32_t>(sp.d);
        s.k = static_cast<uint32_t>(sp.k);
        s.row_begin = row;
        s.prep_begin = prep;
        s.cap = cap;
        TuneEntry tuned{};
        bool have_tuned = false;
        const bool pass = plan.operand != kOperandNone &&
                          sp.k <= static_cast<size_t>(single ? kMaxPassK : kMaxBatchK) &&
                          sp.d <= static_cast<size_t>(single && plan.operand == kOperandF32 ? kMaxPassD : kMaxPackedD);
        s.on_pass = pass ? 1u : 0u;
        if (pass) {
            // The local-origin pass's bound is looser for rows far from their work item's
            // origin: it keeps 64 candidates so those rows still certify.
            s.kp = plan.operand == kOperandF16Local ? 64u : pass_kp(s.k);
            // A row's buffer is four lane segments of cap/4 slots; a compaction leaves kp/4
            // keys in each and a tile appends at most 16 per segment.
            if (s.cap < 4 * s.kp) s.cap = 4 * s.kp;
            if (plan.operand == kOperandF16Local) {
                s.nch = static_cast<uint32_t>(local_chunks(static_cast<int>(sp.d)));
                s.group = 0;
                s.local = 1;
                s.local_c = static_cast<float>(local_product_coeff(static_cast<int>(s.nch)) * (1.0 + 0x1p-20));
                s.local_c2 = static_cast<float>(local_abs_coeff(static_cast<int>(sp.d)) * (1.0 + 0x1p-20));
            } else if (plan.operand == kOperandBf16) {
                // Single calls use the compiled chunk counts; the batched kernel takes any.
                s.nch = single ? bf16_chunks(static_cast<int>(sp.d))
                               : static_cast<uint32_t>((3 * sp.d + 3 + 31) / 32);
                s.group = 0;
                // c = h + l + r with |r| <= 2^-16 |c|: dropping l.l and the residuals costs
                // <= 3.02 * 2^-16 |a||b|; |b|^2 = nh + nl + nl2 to 2^-24; products of bf16 are
                // exact in f32, the accumulation is bounded as 2 * (terms + 1) * 2^-24.
                s.split_rel = 3.02 * 0x1p-16;
                s.acc_rel = 2.0 * (32.0 * s.nch + 1.0) * 0x1p-24 * 1.05;
                s.norm_rel = 0x1p-24 * 1.01;
            } else {
                s.nch = single ? fp32_chunks(static_cast<int>(sp.d))
                               : static_cast<uint32_t>((sp.d + 1 + 3) / 4);
                s.group = fp32_group(static_cast<int>(s.nch));
                s.split_rel = 0.0;
                s.acc_rel = 2.0 * (4.0 * s.nch + 1.0) * 0x1p-24 * 1.05;
                s.norm_rel = 0.0;
            }
            if (s.nch == 0) return fail(SCRUST_HIP_KERNEL, "HIP kernel knn: no pass variant for d = " + std::to_string(sp.d));
            // Work shape: forced (the tuner), from the tuning table (automatic tile), or the
            // defaults below.
            if (force && force->cap) {
                tuned = *force;
                have_tuned = true;
            } else if (!force && single && (!options || options->tile == 0)) {
                have_tuned = tune_lookup(tune_key(precision, metric, s.kp, sp.n, sp.d), &tuned);
            }
            if (have_tuned) {
                if (tuned.cap != 128 && tuned.cap != 256 && tuned.cap != 512) have_tuned = false;
                else if (tuned.cap >= 4 * s.kp) s.cap = tuned.cap;
            }
            plan.tuned |= have_tuned && !force;
            plan.any_pass = true;
            plan.kps |= s.kp;
            plan.nch_max = std::max(plan.nch_max, static_cast<int>(s.nch));
            plan.pass_flops += 2ull * sp.n * sp.n * sp.d;
        } else {
            for (uint32_t r = 0; r < s.n; r++) plan.exact_rows.push_back(row + r);
        }
        // Arena offsets are stored in the pointer fields until plan_place.
        auto take = [&](size_t bytes) {
            const size_t off = arena;
            arena += align_up(bytes);
            return reinterpret_cast<void*>(off);
        };
        s.coords = static_cast<float*>(take(sp.n * sp.d * sizeof(float)));
        s.norms = static_cast<float*>(take(sp.n * sizeof(float)));
        s.mean = static_cast<double*>(take(sp.d * sizeof(double)));
        s.acc = static_cast<unsigned long long*>(take(sp.d * kAccWordsHost * sizeof(unsigned long long)));
        s.max_norm_bits = static_cast<unsigned*>(take(sizeof(unsigned)));
        // Reference operand rows: n rounded up to whole pairs of 64-row tiles, the padding
        // rows being sentinels (norm feature +inf).
        const size_t ref_rows = (sp.n + 2 * kTileRefsHost - 1) / (2 * kTileRefsHost) * (2 * kTileRefsHost);
        if (pass) {
            // The local-origin pass takes kLocalRows query rows per work item.
            const size_t unit = s.local ? kLocalRows : kPassRows;
            const size_t npad = (sp.n + unit - 1) / unit * unit;
            const uint32_t blocks = static_cast<uint32_t>(npad / kPassRows);
            const uint32_t pairs = static_cast<uint32_t>(ref_rows / (2 * kTileRefsHost));
            const uint32_t set = static_cast<uint32_t>(plan.sets.size());
            // Reference ranges per block: a row's ranges feed one final selection of at
            // most kSelectKeysMax keys, so splits x cap <= 1024.
            const uint32_t max_splits = kSelectKeysMax / s.cap;
            uint32_t splits;
            // A build of more blocks than workgroup slots, or a bf16 build of at least
            // kOrderMinCells cells, scans every block in one range on the spatial order, and the
            // last round's blocks in ranges (below): the order keeps the append bounds short,
            // and on the persistent pass its ball bound skips the far pairs. Ranges from the
            // start cost that order (the cost model gave such builds 2 to 4 ranges of the input
            // order): pass at 80,000 cells 5.14 -> 3.67 ms, at 140,000 17.7 -> 6.9 ms, at
            // 160,000 x 100 34.3 -> 16.4 ms, at 300,000 x 50 fp32 88.6 -> 74.3 ms; serving at
            // 0.8 of saturation p99 62-65 -> 26-27 ms, saturation 370 -> 426-449 requests/s. The
            // fp32 pass of a single round keeps the model's ranges (80,000 cells: 9.04 against
            // 9.84 ms in one range).
            const bool one_range = single && !s.local &&
                                   (blocks > slots || (plan.operand == kOperandBf16 && sp.n >= kOrderMinCells));
            // An item's fixed cost is about 64k cycles; a tile pair costs 2 x 4 x 4 x nch
            // MFMAs of 16 (bf16) or 32 (fp32) cycles.
            const double pair_cycles = 32.0 * s.nch * (plan.operand == kOperandBf16 ? 16.0 : 32.0);
            const double overhead = 65536.0 / pair_cycles;
            if (s.local) {
                splits = 1;  // every work item scans all references, from its group's tiles
                s.chunk_pairs = pairs;
            } else if (single) {
                splits = one_range ? 1u : choose_splits(blocks, pairs, slots, max_splits, overhead);
                if (have_tuned && tuned.splits >= 1 && tuned.splits <= max_splits) {
                    // A tuned count is taken when it cuts the scan into exactly that many
                    // whole-pair ranges.
                    const uint32_t len = (pairs + tuned.splits - 1) / tuned.splits;
                    if ((pairs + len - 1) / len == tuned.splits) splits = tuned.splits;
                }
                s.chunk_pairs = (pairs + splits - 1) / splits;
            } else {
                const uint32_t pieces = static_cast<uint32_t>(std::min<uint64_t>(max_splits, (pairs + batch_len - 1) / batch_len));
                splits = partition_even(set, blocks, pairs, pieces, plan.items);
                s.chunk_pairs = (pairs + splits - 1) / splits;
            }
            s.npad = static_cast<uint32_t>(npad);
            s.splits = splits;
            // Spatial order (single datasets): automatic from kOrderMinCells cells when a
            // block scans all references in one range (the scan then starts at the block's
            // own part of the order; a range that starts elsewhere meets the cells in sorted
            // order, which appends more than a random order), or as the options ask.
            uint32_t order_knob = options ? options->order : 0u;
            if (force && force->cap) order_knob = force->order;
            else if (have_tuned && order_knob == 0u && (tuned.order == 1u || tuned.order == 2u)) order_knob = tuned.order;
            if (s.local || (single && (order_knob == 2u || (order_knob == 0u && sp.n >= kOrderMinCells && splits == 1)))) {
                plan.ordered = true;
                plan.order = order_layout(sp.n, sp.d);
                s.perm = static_cast<const uint32_t*>(take(sp.n * sizeof(uint32_t)));
                s.inv_perm = static_cast<const uint32_t*>(take(sp.n * sizeof(uint32_t)));
                plan.order_offset = reinterpret_cast<size_t>(take(plan.order.total));
                s.rotate = splits == 1 ? 1u : 0u;
            }
            // The launch's last round (single datasets scanned in one range): when it would
            // fill more than half of the workgroup slots, some CUs run two of its workgroups
            // and the others one, and the round lasts as long as a full one. The last
            // blocks % slots blocks then scan in kTailSplits ranges (of the rotated sequence in
            // spatial order), which spreads that round's work over every slot.
            const uint64_t rem = blocks % slots;
            const uint32_t tail_splits = std::min(kTailSplits, max_splits);
            // Such builds split any last round: the ranges go to the workgroups that free up
            // first, where whole blocks would leave the others idle.
            if (single && !s.local && splits == 1 && blocks > slots && rem > 0 && (rem > slots / 2 || one_range) &&
                tail_splits > 1) {
                const uint32_t len = (pairs + tail_splits - 1) / tail_splits;
                if ((pairs + len - 1) / len == tail_splits) {
                    splits = tail_splits;
                    s.splits = splits;
                    s.chunk_pairs = len;
                    s.tail_row0 = static_cast<uint32_t>((blocks - rem) * kPassRows);
                    s.select_e1 = next_pow2(std::max<uint32_t>(2, (s.cap + 63) / 64));
                }
            }
            s.select_e = next_pow2(std::max<uint32_t>(2, (splits * s.cap + 63) / 64));
            // Persistent pass: a single dataset on split tiles in pairs whose blocks outnumber
            // the workgroup slots. Its candidate buffers are per slot (kPassRows rows each)
            // instead of per row, the pass selects each whole block's candidates itself, and
            // rows scanned in ranges keep each range's top kp for merge_kernel. Not on a
            // shared GPU (server requests): its workgroups hold every slot until the launch
            // ends, where the other requests' kernels would wait for them.
            const bool persist = single && !shared && !s.local && split_pass_nch(plan.operand, s.nch) &&
                                 s.cap >= 256 && s.cap <= static_cast<uint32_t>(kCapMax) && blocks > slots;
            if (persist) {
                s.persist = 1;
                s.slots = static_cast<uint32_t>(slots);
                s.select_e = 0;
                s.select_e1 = 0;
            } else {
                plan.select_widths |= s.select_e | s.select_e1;
            }
            if (s.local) {
                s.pack_src = static_cast<float*>(take(ref_rows * 128ull * s.nch));
                s.pack_query = static_cast<uint16_t*>(take(sp.n * 64ull * s.nch));
                s.origins = static_cast<float*>(take(npad / kLocalRows * 128ull * s.nch));
                s.anorm = static_cast<float*>(take(sp.n * sizeof(float)));
            } else if (plan.operand == kOperandBf16) {
                const size_t rowb = 64ull * s.nch;
                s.pack_ref = static_cast<uint16_t*>(take(ref_rows * rowb));
                // The persistent pass forms its query fragments from the coordinates.
                if (!s.persist) s.pack_query = static_cast<uint16_t*>(take(sp.n * rowb));
            } else {
                s.pack_f32 = static_cast<float*>(take(ref_rows * 16ull * s.group));
            }
            if (s.persist) {
                s.buf = static_cast<uint64_t*>(take(static_cast<size_t>(s.slots) * kPassRows * s.cap * sizeof(uint64_t)));

                if (s.splits > 1)
                    s.tail_cand = static_cast<uint64_t*>(
                        take(static_cast<size_t>(s.n - s.tail_row0) * s.splits * s.kp * sizeof(uint64_t)));
                s.seg_counts = reinterpret_cast<uint32_t*>(count_words * sizeof(uint32_t));  // unused
            } else {
                s.buf = static_cast<uint64_t*>(take(range_rows(s) * s.cap * sizeof(uint64_t)));
                s.seg_counts = reinterpret_cast<uint32_t*>(count_words * sizeof(uint32_t));  // region offset
                count_words += range_rows(s);
            }
            // Bounding balls of the tile pairs and query blocks (the persistent pass on the
            // spatial order): there a block and a pair of one part of the order are compact,
            // and a pair out of reach of every row of the block is skipped. Builds of up to 512
            // blocks keep none: at 100,000 cells the bound skipped 4% of the pairs and its
            // tests cost 10% of the pass.
            const bool wide_rounds = single && plan.operand == kOperandBf16 && !s.local && s.nch > 5 &&
                                     bf16_coarse_chunks(static_cast<int>(s.nch)) < static_cast<int>(s.nch) &&
                                     blocks > slots && s.cap <= static_cast<uint32_t>(kCapMax);
            // The fp32 pass of 129-256 features (the path of every precision there) keeps
            // them too; at up to 128 it stays the plain fp32 kernel.
            const bool wide_f32 = single && plan.operand == kOperandF32 && sp.d > static_cast<size_t>(kMaxPackedD) &&
                                  blocks > slots && s.cap <= static_cast<uint32_t>(kCapMax);
            if ((s.persist || wide_rounds || wide_f32) && plan.ordered && sp.d <= kBallMaxDims) {
                s.ball_stride = (static_cast<uint32_t>(sp.d) + 3u) / 4u * 4u;
                s.balls = static_cast<float*>(take((pairs + blocks) * (s.ball_stride + 1ull) * sizeof(float)));
            }
            s.cand = static_cast<uint64_t*>(take(sp.n * s.kp * sizeof(uint64_t)));
            s.theta = static_cast<float*>(take(sp.n * sizeof(float)));
        }
        plan.sets.push_back(s);
        plan.shape.max_row_chunks = std::max<uint32_t>(plan.shape.max_row_chunks, static_cast<uint32_t>((sp.n + colsum_rows(s.n) - 1) / colsum_rows(s.n)));
        plan.shape.max_col_groups = std::max<uint32_t>(plan.shape.max_col_groups, static_cast<uint32_t>((sp.d + 63) / 64));
        plan.shape.max_d = std::max<uint32_t>(plan.shape.max_d, static_cast<uint32_t>(sp.d));
        if (pass && (plan.operand == kOperandBf16 || plan.operand == kOperandF16Local))
            plan.shape.max_prep_blocks = std::max<uint32_t>(plan.shape.max_prep_blocks,
                                                            static_cast<uint32_t>((ref_rows + 63) / 64));
        else
            plan.shape.direct_sets++;
        if (pass && s.local)
            plan.shape.max_local_blocks = std::max<uint32_t>(plan.shape.max_local_blocks, s.npad / kLocalRows);
        row += s.n;
        prep += pass ? static_cast<uint32_t>(ref_rows) : s.n;
    }
    plan.shape.total_rows = row;
    plan.shape.total_prep = prep;
    // Workspace of the split scan that recomputes a short list of uncertified rows.
    for (const DatasetDesc& s : plan.sets)
        if (s.on_pass) plan.split_ranges = std::max<uint32_t>(plan.split_ranges, (s.n + kSplitRange - 1) / kSplitRange);
    if (plan.split_ranges) {
        plan.split_offset = arena;
        arena += align_up(static_cast<size_t>(kSplitRowsMax) * plan.split_ranges * kSplitKeep * sizeof(uint64_t));
    }
    // The segment counts of all datasets: one region, cleared before every pass.
    plan.count_words = count_words;
    plan.counts_offset = arena;
    arena += align_up(count_words * sizeof(uint32_t) + 1);
    for (DatasetDesc& s : plan.sets)
        if (s.on_pass)
            s.seg_counts = reinterpret_cast<uint32_t*>(plan.counts_offset + reinterpret_cast<size_t>(s.seg_counts));
    if (plan.any_pass && !single) {
        // One persistent launch per compiled chunk class (narrow rows keep the occupancy
        // of their own variant), each over its items largest first: an item costs about
        // its tile count (the references it scans).
        auto cls = [&](const PassItem& it) { return batch_chunks(plan.operand, static_cast<int>(plan.sets[it.set].nch)); };
        for (const PassItem& it : plan.items)
            if (cls(it) == 0) return fail(SCRUST_HIP_KERNEL, "HIP kernel knn: no batched pass variant");
        std::stable_sort(plan.items.begin(), plan.items.end(), [&](const PassItem& a, const PassItem& b) {
            const int ca = cls(a), cb = cls(b);
            if (ca != cb) return ca < cb;
            return a.tile_end - a.tile_begin > b.tile_end - b.tile_begin;
        });
        for (size_t i = 0; i < plan.items.size(); i++) {
            const int c = cls(plan.items[i]);
            if (plan.groups.empty() || plan.groups.back().nch_max != c)
                plan.groups.push_back(PassGroup{c, static_cast<uint32_t>(i), 0});
            plan.groups.back().count++;
        }
        if (plan.groups.size() > static_cast<size_t>(kScalarCount - kScalarItemCounter))
            return fail(SCRUST_HIP_KERNEL, "HIP kernel knn: too many batched pass groups");
    }
    if (!plan.any_pass) plan.operand = kOperandNone;
    plan.shape.operand = plan.operand;
    plan.arena_bytes = arena;
    plan.meta_bytes = align_up(plan.sets.size() * sizeof(DatasetDesc)) +
                      align_up(plan.items.size() * sizeof(PassItem) + 1) +
                      align_up(static_cast<size_t>(row) * sizeof(uint32_t) + 1) +
                      align_up(kScalarCount * sizeof(uint32_t));
    return SCRUST_HIP_OK;
}

scrust_hip_status plan_place(Plan& plan, void* arena, void* meta, hipStream_t stream,
                             void* pinned) {
    char* base = static_cast<char*>(arena);
    plan.arena_base = arena;
    // Every field passed here holds an arena offset (possibly 0) from plan_layout.
    auto fix = [&](auto*& p) {
        using T = std::remove_reference_t<decltype(*p)>;
        p = reinterpret_cast<T*>(base + reinterpret_cast<size_t>(p));
    };
    for (DatasetDesc& s : plan.sets) {
        fix(s.coords);
        fix(s.norms);
        fix(s.mean);
        fix(s.acc);
        fix(s.max_norm_bits);
        if (s.on_pass) {
            if (s.local) {
                fix(s.pack_src);
                fix(s.pack_query);
                fix(s.origins);
                fix(s.anorm);
            } else if (plan.operand == kOperandBf16) {
                fix(s.pack_ref);
                if (!s.persist) fix(s.pack_query);
            } else {
                fix(s.pack_f32);
            }
            fix(s.buf);
            fix(s.seg_counts);
            if (s.persist && s.splits > 1) fix(s.tail_cand);
            if (s.balls) fix(s.balls);
            fix(s.cand);
            fix(s.theta);
            if (s.perm) {
                fix(s.perm);
                fix(s.inv_perm);
            }
        }
    }
    char* m = static_cast<char*>(meta);
    plan.d_sets = reinterpret_cast<DatasetDesc*>(m);
    m += align_up(plan.sets.size() * sizeof(DatasetDesc));
    plan.d_items = reinterpret_cast<PassItem*>(m);
    m += align_up(plan.items.size() * sizeof(PassItem) + 1);
    plan.d_list = reinterpret_cast<uint32_t*>(m);
    m += align_up(static_cast<size_t>(plan.shape.total_rows) * sizeof(uint32_t) + 1);
    plan.d_scalars = reinterpret_cast<uint32_t*>(m);
    // The persistent pass's work queues: scalars cleared before every pass.
    static_assert(kScalarItemCounter + kQueueWords <= kScalarCount, "queue words among the scalars");
    for (DatasetDesc& s : plan.sets)
        if (s.persist || s.balls) s.queue = plan.d_scalars + kScalarItemCounter;
    const size_t sets_bytes = plan.sets.size() * sizeof(DatasetDesc);
    const size_t items_bytes = plan.items.size() * sizeof(PassItem);
    const size_t list_bytes = plan.exact_rows.size() * sizeof(uint32_t);
    hipError_t err = hipSuccess;
    if (pinned) {
        // Stage the metadata in pinned memory with the device layout: one async copy.
        char* h = static_cast<char*>(pinned);
        const size_t items_at = reinterpret_cast<char*>(plan.d_items) - static_cast<char*>(meta);
        const size_t list_at = reinterpret_cast<char*>(plan.d_list) - static_cast<char*>(meta);
        std::memcpy(h, plan.sets.data(), sets_bytes);
        if (items_bytes) std::memcpy(h + items_at, plan.items.data(), items_bytes);
        if (list_bytes) std::memcpy(h + list_at, plan.exact_rows.data(), list_bytes);
        err = hipMemcpyAsync(meta, h, list_at + list_bytes, hipMemcpyHostToDevice, stream);
    } else {
        err = hipMemcpyAsync(plan.d_sets, plan.sets.data(), sets_bytes, hipMemcpyHostToDevice, stream);
        if (err == hipSuccess && items_bytes)
            err = hipMemcpyAsync(plan.d_items, plan.items.data(), items_bytes, hipMemcpyHostToDevice, stream);
        if (err == hipSuccess && list_bytes)
            err = hipMemcpyAsync(plan.d_list, plan.exact_rows.data(), list_bytes, hipMemcpyHostToDevice, stream);
    }
    if (err != hipSuccess) return hip_failure(err, "upload launch metadata");
    return SCRUST_HIP_OK;
}

scrust_hip_status plan_enqueue(const Plan& plan, bool single, int compute_units,
                               hipStream_t stream, hipEvent_t* events, hipStream_t side,
                               hipEvent_t fork, hipEvent_t join) {
    auto mark = [&](int i) -> hipError_t { return events ? hipEventRecord(events[i], stream) : hipSuccess; };
    hipError_t err = mark(0);
    if (err != hipSuccess) return hip_failure(err, "record event");
    // Scalars: flags = 0, exact-list count = the exact-path rows already listed, item
    // counter = 0. Cleared with a kernel (graph-safe).
    launch_fill_u32(plan.d_scalars, kScalarCount, 0u, stream);
    launch_fill_u32(plan.d_scalars + kScalarListCount, 1, static_cast<uint32_t>(plan.exact_rows.size()), stream);
    unsigned* flags = plan.d_scalars + kScalarFlags;
    if (plan.count_words)
        launch_fill_u32(reinterpret_cast<uint32_t*>(static_cast<char*>(plan.arena_base) + plan.counts_offset),
                        plan.count_words, 0u, stream);
    if (plan.ordered) {
        const DatasetDesc& s = plan.sets[0];
        err = order_enqueue(s.x, s.n, s.d, static_cast<char*>(plan.arena_base) + plan.order_offset, plan.order,
                            const_cast<uint32_t*>(s.perm), const_cast<uint32_t*>(s.inv_perm), stream);
        if (err != hipSuccess) return hip_failure(err, "launch row order");
    }
    launch_prepare_all(plan.d_sets, plan.shape, flags, stream);
    if ((err = hipGetLastError()) != hipSuccess) return hip_failure(err, "launch prepare");
    if ((err = mark(1)) != hipSuccess) return hip_failure(err, "record event");
    if (plan.any_pass) {
        bool ok;
        if (single) {
            if (plan.sets[0].balls) launch_balls(plan.d_sets, plan.sets[0], stream);
            ok = launch_pass(plan.d_sets, plan.sets[0], plan.operand, stream);
        } else {
            ok = true;
            // With a side stream, every class but the first runs on it, in parallel with the
            // first: a class's persistent workgroups take the slots the other's tail frees.
            const bool fan = side && fork && join && plan.groups.size() > 1;
            if (fan) {
                if ((err = hipEventRecord(fork, stream)) != hipSuccess) return hip_failure(err, "record event");
                if ((err = hipStreamWaitEvent(side, fork, 0)) != hipSuccess) return hip_failure(err, "wait event");
            }
            for (size_t g = plan.groups.size(); g-- > 0 && ok;) {
                const PassGroup& pg = plan.groups[g];
                const unsigned blocks = static_cast<unsigned>(std::min<size_t>(pg.count, 2ull * compute_units));
                ok = launch_pass_batch(plan.d_sets, plan.d_items + pg.begin, pg.count,
                                       plan.d_scalars + kScalarItemCounter + g, plan.operand,
                                       pg.nch_max, blocks, fan && g > 0 ? side : stream);
            }
            if (fan) {
                if ((err = hipEventRecord(join, side)) != hipSuccess) return hip_failure(err, "record event");
                if ((err = hipStreamWaitEvent(stream, join, 0)) != hipSuccess) return hip_failure(err, "wait event");
            }
        }
        if (!ok) return fail(SCRUST_HIP_KERNEL, "HIP kernel knn: no compiled pass variant for this shape");
        if ((err = hipGetLastError()) != hipSuccess) return hip_failure(err, "launch candidate pass");
    }
    if ((err = mark(2)) != hipSuccess) return hip_failure(err, "record event");
    if (plan.any_pass) {
        if (single) launch_merge(plan.d_sets, plan.sets[0], stream);
        launch_select(plan.d_sets, plan.shape, plan.select_widths, stream);
        if ((err = hipGetLastError()) != hipSuccess) return hip_failure(err, "launch candidate selection");
        launch_rerank(plan.d_sets, plan.shape, plan.kps, plan.d_list,
                      plan.d_scalars + kScalarListCount, stream);
        if ((err = hipGetLastError()) != hipSuccess) return hip_failure(err, "launch rerank");
    }
    if ((err = mark(3)) != hipSuccess) return hip_failure(err, "record event");
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>(plan.shape.total_rows, 4096));
    uint64_t* split = plan.split_ranges ? reinterpret_cast<uint64_t*>(static_cast<char*>(plan.arena_base) + plan.split_offset)
                                        : nullptr;
    launch_exact(plan.d_sets, plan.shape, plan.d_list, plan.d_scalars + kScalarListCount, blocks, flags, split, plan.split_ranges,
                 stream);
    if ((err = hipGetLastError()) != hipSuccess) return hip_failure(err, "launch exact rows");
    if ((err = mark(4)) != hipSuccess) return hip_failure(err, "record event");
    return SCRUST_HIP_OK;
}

}  // namespace scrust_hip
