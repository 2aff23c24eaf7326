// SPDX-License-Identifier: 0BSD
//
// The plan compiler (host only): words -> rows -> sorted rows -> instructions. See
// plan_compiler.hpp for the stages.
#include "plan_compiler.hpp"

#include <algorithm>

namespace pacuda {

namespace {

// A field's width: one of the plan's six widths (W_P .. W_K) or a fixed count of bits.
enum : uint8_t { FP = W_P, FA = W_A, FF = W_F, FX = W_X, FS = W_S, FK = W_K, FIXED = 0x80 };
constexpr uint8_t bits(int n) { return (uint8_t)(FIXED | n); }

struct Layout {
    uint8_t fields;
    uint8_t width[8];
};

// The fields of each op, in word order after the 4-bit op.
//   1 clear                    base P, count P
//   2-6 contributions          init 1, scale slot S, parent 0 P, parent 1 P, destination P,
//                              form 0 F, form 1 F, factor X
//   7-9 propagators            base P, form F, factor X
//   10 identity                base P, count K, factor X
//   11 closure                 parent 0 P, parent 1 P, amplitude A, factor X, first component
//                              factor X, count K
//   12 source                  destination P, form F, factor X, mass slot S, template 10
//   13 symjit row              init 1, parent 0 P, parent 1 P, destination P, form 0 F,
//                              form 1 F, factor X
//   14 binding                 kernel 16, inputs 8, output pairs 8, direct 1; then its codes,
//                              two u32 per data word, low half first
constexpr Layout CONTRIBUTION = {8, {bits(1), FS, FP, FP, FP, FF, FF, FX}};
constexpr Layout PROPAGATOR = {3, {FP, FF, FX}};
constexpr Layout LAYOUTS[16] = {
    {0, {}},                                         // 0: not an op
    {2, {FP, FP}},                                   // 1
    CONTRIBUTION, CONTRIBUTION, CONTRIBUTION, CONTRIBUTION, CONTRIBUTION,  // 2-6
    PROPAGATOR, PROPAGATOR, PROPAGATOR,              // 7-9
    {3, {FP, FK, FX}},                               // 10
    {6, {FP, FP, FA, FX, FX, FK}},                   // 11
    {5, {FP, FF, FX, FS, bits(10)}},                 // 12
    {7, {bits(1), FP, FP, FP, FF, FF, FX}},          // 13
    {4, {bits(16), bits(8), bits(8), bits(1)}},      // 14
    {0, {}},                                         // 15: not an op
};

bool is_contribution(int op) { return op >= 2 && op <= 6; }

}  // namespace

ContributionShape contribution_shape(int op) {
    // op 2/3: Weyl x vector -> Weyl; op 4: tensor x vector -> vector; op 5: vector x vector ->
    // tensor; op 6: vector x vector -> vector
    switch (op) {
    case 4: return {6, 4, 4};
    case 5: return {4, 4, 6};
    case 6: return {4, 4, 4};
    default: return {2, 4, 2};
    }
}

RowSet decode_rows(const uint64_t *words, uint32_t count, const uint8_t *width) {
    RowSet set;
    set.rows.reserve(count);
    for (int k = 0; k < 6; k++)
        if (width[k] > 32) {
            set.error = "a field width above 32 bits";
            return set;
        }
    uint32_t n = 0;
    while (n < count) {
        uint64_t word = words[n++];
        PlanRow row;
        row.op = (uint8_t)(word & 15);
        word >>= 4;
        const Layout &layout = LAYOUTS[row.op];
        for (int k = 0; k < layout.fields; k++) {
            const uint8_t w = layout.width[k];
            const int b = (w & FIXED) ? (w & 0x7f) : width[w];
            row.f[k] = b ? (uint32_t)(word & ((1ull << b) - 1)) : 0u;
            word = b < 64 ? word >> b : 0;
        }
        if (row.op == 14) {
            // the binding's input codes then its output components; data words the plan does not
            // have read as zeros
            row.data_at = (uint32_t)set.data.size();
            row.data_len = row.f[1] + row.f[2];
            for (uint32_t k = 0; k < row.data_len; k += 2) {
                const uint64_t d = n < count ? words[n] : 0;
                n++;
                set.data.push_back((uint32_t)d);
                if (k + 1 < row.data_len) set.data.push_back((uint32_t)(d >> 32));
            }
        }
        set.rows.push_back(row);
    }
    return set;
}

void sort_runs(std::vector<PlanRow> &rows) {
    struct Span {
        uint32_t start, end;
    };
    std::vector<Span> dst;
    size_t first = 0;
    while (first < rows.size()) {
        if (!is_contribution(rows[first].op)) {
            first++;
            continue;
        }
        size_t last = first;
        while (last < rows.size() && is_contribution(rows[last].op)) last++;
        // the destinations of the run as spans [base, base + components), one per base (its
        // widest), ordered by base; they must not overlap
        dst.clear();
        for (size_t r = first; r < last; r++) dst.push_back({rows[r].f[4], rows[r].f[4] + (uint32_t)contribution_shape(rows[r].op).cd});
        std::sort(dst.begin(), dst.end(), [](const Span &a, const Span &b) {
            return a.start < b.start || (a.start == b.start && a.end > b.end);
        });
        size_t kept = 0;
        bool safe = true;
        for (size_t k = 0; k < dst.size(); k++) {
            if (kept && dst[kept - 1].start == dst[k].start) continue;  // the same base: its widest span is first
            if (kept && dst[k].start < dst[kept - 1].end) safe = false;
            dst[kept++] = dst[k];
        }
        dst.resize(kept);
        // no parent of the run may touch a destination of the run
        auto touches = [&](uint32_t start, uint32_t components) {
            const uint32_t end = start + components;
            auto it = std::upper_bound(dst.begin(), dst.end(), end, [](uint32_t v, const Span &s) { return v <= s.start; });
            return it != dst.begin() && std::prev(it)->end > start;
        };
        for (size_t r = first; r < last && safe; r++) {
            const ContributionShape s = contribution_shape(rows[r].op);
            safe = !touches(rows[r].f[2], (uint32_t)s.c0) && !touches(rows[r].f[3], (uint32_t)s.c1);
        }
        if (safe)
            std::stable_sort(rows.begin() + first, rows.begin() + last,
                             [](const PlanRow &a, const PlanRow &b) { return a.f[4] < b.f[4]; });
        first = last;
    }
}

namespace {

// The instruction of one row (not a binding): the kind, op and flags in w[0], the fields after.
Instr instruction(const PlanRow &r, uint32_t flags, uint32_t binding) {
    Instr o{};
    int kind;
    if (is_contribution(r.op)) {
        kind = K_CONTRIB;
        if (r.f[0]) flags |= F_INIT;
        for (int k = 1; k < 8; k++) o.w[k] = r.f[k];
    } else if (r.op == 13) {
        kind = K_SYMJIT;
        if (r.f[0]) flags |= F_INIT;
        o.w[1] = r.f[1];
        o.w[2] = r.f[2];
        o.w[3] = r.f[3];
        o.w[4] = r.f[4] | r.f[5] << 16;
        o.w[5] = r.f[6];
        o.w[6] = binding;
    } else {
        static const int KIND[13] = {0, K_CLEAR, 0, 0, 0, 0, 0, K_PROP, K_PROP, K_PROP, K_IDENTITY, K_CLOSURE, K_SOURCE};
        kind = KIND[r.op];
        for (int k = 0; k < 7; k++) o.w[k + 1] = r.f[k];
    }
    o.w[0] = (uint32_t)kind | (uint32_t)r.op << 8 | flags << 16;
    return o;
}

// The extent of every table an instruction indexes (bindings checked first).
struct Extents {
    uint64_t need[N_COUNT] = {0, 0, 0, 0, 0};
    void use(int table, uint64_t first, uint64_t count) {
        if (count) need[table] = std::max<uint64_t>(need[table], first + count);
    }
    void add(const Instr &o, const std::vector<uint32_t> &bindings) {
        const int kind = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff;
        switch (kind) {
        case K_CONTRIB: {
            const ContributionShape s = contribution_shape(op);
            use(N_CURRENT, o.w[2], s.c0);
            use(N_CURRENT, o.w[3], s.c1);
            use(N_CURRENT, o.w[4], s.cd);
            use(N_SCALE, o.w[1], 1);
            use(N_FACTOR, o.w[7], 1);
            if (op == 6) {
                use(N_FORM, o.w[5], 1);
                use(N_FORM, o.w[6], 1);
            }
            break;
        }
        case K_CLEAR: use(N_CURRENT, o.w[1], o.w[2]); break;
        case K_PROP:
            use(N_CURRENT, o.w[1], op == 9 ? 4 : 2);
            use(N_FORM, o.w[2], 1);
            use(N_FACTOR, o.w[3], 1);
            break;
        case K_IDENTITY:
            use(N_CURRENT, o.w[1], o.w[2]);
            use(N_FACTOR, o.w[3], 1);
            break;
        case K_CLOSURE:
            use(N_CURRENT, o.w[1], o.w[6]);
            use(N_CURRENT, o.w[2], o.w[6]);
            use(N_AMPLITUDE, o.w[3], 1);
            use(N_FACTOR, o.w[4], 1);
            use(N_FACTOR, o.w[5], o.w[6]);
            break;
        case K_SOURCE: {
            const uint32_t family = o.w[5] & 7;
            use(N_CURRENT, o.w[1], family == 0 ? 1 : family == 1 ? 2 : 4);
            use(N_FORM, o.w[2], 1);
            use(N_FACTOR, o.w[3], 1);
            use(N_SCALE, o.w[4], 1);
            break;
        }
        case K_SYMJIT: {
            const uint32_t *b = bindings.data() + o.w[6];
            const uint32_t inputs = b[1], pairs = b[2];
            const uint32_t parent[2] = {o.w[1], o.w[2]}, form[2] = {o.w[4] & 0xffff, o.w[4] >> 16};
            use(N_FACTOR, o.w[5], 1);
            for (uint32_t k = 0; k < inputs; k++) {
                const uint32_t code = b[4 + k], which = (code >> 3) & 1, index = code >> 5;
                switch (code & 7) {
                case SB_PARENT: use(N_CURRENT, (uint64_t)parent[which] + index, 1); break;
                case SB_MOMENTUM: use(N_FORM, ((uint64_t)form[which] * 4 + index) / 4, 1); break;
                case SB_DESTINATION: use(N_CURRENT, (uint64_t)o.w[3] + index, 1); break;
                case SB_SCALE: use(N_SCALE, index, 1); break;
                default: break;  // the factor
                }
            }
            for (uint32_t j = 0; j < pairs; j++) use(N_CURRENT, (uint64_t)o.w[3] + b[4 + inputs + j], 1);
            break;
        }
        default: break;
        }
    }
};

}  // namespace

CompiledProgram compile_program(const uint64_t *words, uint32_t count, const uint8_t *width, bool sort,
                                const KernelLimits &kernels) {
    CompiledProgram p;
    RowSet set = decode_rows(words, count, width);
    if (set.error) {
        p.error = set.error;
        return p;
    }
    for (const PlanRow &r : set.rows) {
        if (r.op == 0 || r.op == 15) {
            p.error = "a word with an unknown op";
            return p;
        }
        if ((r.op == 13 || r.op == 14) && kernels.count == 0) {
            p.error = "a symjit row, but this runner has no symjit kernels";
            return p;
        }
    }
    if (sort) sort_runs(set.rows);
    const std::vector<PlanRow> &rows = set.rows;
    uint32_t binding = UINT32_MAX;  // the binding the next symjit row uses: the last one before it
    for (size_t n = 0; n < rows.size(); n++) {
        const PlanRow &r = rows[n];
        if (r.op == 14) {
            binding = (uint32_t)p.bindings.size();
            p.bindings.insert(p.bindings.end(), {r.f[0], r.f[1], r.f[2], r.f[3]});
            p.bindings.insert(p.bindings.end(), set.data.begin() + r.data_at, set.data.begin() + r.data_at + r.data_len);
            continue;
        }
        uint32_t flags = 0;
        if (is_contribution(r.op)) {
            // a destination's rows are adjacent after the sort; the first loads or starts the sum,
            // the last stores it
            const bool same_before = n > 0 && is_contribution(rows[n - 1].op) && rows[n - 1].f[4] == r.f[4];
            const bool same_after = n + 1 < rows.size() && is_contribution(rows[n + 1].op) && rows[n + 1].f[4] == r.f[4];
            flags = (same_before ? 0u : (uint32_t)F_FIRST) | (same_after ? 0u : (uint32_t)F_LAST);
        }
        if (r.op == 13) {
            p.symjit = true;
            // the binding: a kernel the runner has, inputs and outputs within its limits, codes
            // within the table
            const bool ok = binding != UINT32_MAX && p.bindings[binding] < kernels.count &&
                            p.bindings[binding + 1] <= kernels.max_in && 2ull * p.bindings[binding + 2] <= kernels.max_out;
            if (!ok) {
                p.error = "a symjit row without a valid binding";
                return p;
            }
        }
        p.ops.push_back(instruction(r, flags, binding));
    }
    Extents e;
    for (const Instr &o : p.ops) e.add(o, p.bindings);
    std::copy(e.need, e.need + N_COUNT, p.need);
    while (p.sources < p.ops.size() && (p.ops[p.sources].w[0] & 0xff) == K_SOURCE) p.sources++;
    for (size_t n = p.sources; n < p.ops.size(); n++)
        if ((p.ops[n].w[0] & 0xff) == K_SOURCE) {
            p.error = "a source row follows other rows";
            return p;
        }
    return p;
}

}  // namespace pacuda
