// SPDX-License-Identifier: 0BSD
//
// Cooperative programs (host pass of the fused path). The one-wavefront-per-unit kernels keep
// twelve 64-point units in flight per compute unit, so each point gets about a dozen on-chip
// planes and the rest of its live currents (160 to 280 planes for g g > g g g) cross L2 to HBM.
// A cooperative program instead runs one unit on a whole workgroup: its W wavefronts share the
// unit's LDS region (up to 159 planes of 64 lanes per point) and a small slot of its own, and
// each runs a share of the rows. The program's true dependences are wide (g g > g g g: 114 tasks
// in 5 levels), so the rows are cut into phases that end at workgroup barriers:
//
//   - tasks: a contribution run (FIRST to LAST on one destination, its sums in the accumulators)
//     with the in-place rows forwarded from it (F_FROMACC), or a single other row;
//   - dependences: a task reads a current after the task that last wrote it; a task that writes
//     a current comes after its last writer and after every task that read the value it
//     replaces; closures into one amplitude keep their order. Every order that honours these
//     runs the same operations on the same operands and adds the same terms to each destination
//     in the same order, so every value keeps its bits;
//   - phases: each task in the first phase after all its dependences (as soon as possible);
//     within a phase the tasks go to the least-loaded wavefront, longest first;
//   - placement: a current (the plane range a source, a clear or an initialising run creates),
//     an amplitude plane and a momentum form (two planes) each live from the phase that writes
//     them to the last phase that reads them (forms from the forms phase that opens each unit,
//     amplitudes to the reduction that closes it). Objects are ranked by accesses per plane per
//     square root of their lifetime and placed first-fit in LDS over their phases; room freed in
//     a phase is reused only from the next phase on (the barrier between orders the reads before
//     the writes). The rest go to the workgroup's slot the same way, without a limit.
//
// Each wavefront's stream holds its tasks phase by phase with a barrier row after each phase but
// the last, so all wavefronts meet the same barriers, and one stream holds a fetch row that takes
// the workgroup's next unit off the work queue. The forms phase, the barrier after it, the
// barrier before the reduction and the reduction itself are the kernel's; the kernel computes the
// next unit's forms while one wavefront reduces this unit's amplitudes, so an amplitude's LDS room
// is never a form's.
#pragma once
#include <stdint.h>

#include <string.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

#include "device.hpp"
#include "plan_compiler.hpp"
#include "program.hpp"

namespace pacuda {

struct CoopProgram {
    std::vector<Instr> ops;            // wavefront 0's stream, then wavefront 1's, ...
    std::vector<uint32_t> start;       // W + 1 stream offsets
    std::vector<uint32_t> split;       // per wavefront: where its first phase (sources and clears) ends in `ops`
    std::vector<Instr> serial;         // the same rows in one order (phase by phase, wavefront by wavefront), forms and amplitudes
                                       // under their plan indices and LDS planes after the slot's: for the CPU oracle
    std::vector<uint32_t> form_field;  // per form: FORM_LDS | LDS plane, or the slot form
    std::vector<uint32_t> amp_field;   // per amplitude plane: LDS_PLANE + LDS plane, or the slot amplitude
    uint32_t lds_planes = 0, slot_planes = 0, phases = 0, waves = 0;
    uint64_t slot_accesses = 0, lds_accesses = 0;  // current components per point in the slot / in LDS
    bool ok = false;
};

namespace coop_detail {

// a row's rough cost in arithmetic (for balancing the wavefronts of a phase), and the charges per
// current component a row moves through LDS or the slot
constexpr uint32_t COST_LDS = 2, COST_SLOT = 16;
inline uint32_t row_cost(const Instr &o) {
    const int kind = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff;
    switch (kind) {
    case K_CONTRIB: return op == 6 ? 160 : op == 5 ? 100 : op == 4 ? 80 : 40;
    case K_PROP: return op == 9 ? 50 : 30;
    case K_IDENTITY: return 4 * o.w[2] + 4;
    case K_CLOSURE: return 8 * o.w[6] + 4;
    case K_SOURCE: return 60;
    case K_CLEAR: return o.w[2] + 4;
    case K_SYMJIT: return 60;
    default: return 10;
    }
}

struct Field {
    int word;
    uint32_t len;
    bool read, write, create;
};

// the current-plane fields of row o (every word that names a current, with what the row does
// through it); -1 when the row is not one this pass knows
inline int current_fields(const Instr &o, Field *f) {
    const int kind = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff, flags = o.w[0] >> 16;
    switch (kind) {
    case K_SOURCE: {
        const uint32_t family = o.w[5] & 7;
        f[0] = {1, family == 0 ? 1u : family == 1 ? 2u : 4u, false, true, true};
        return 1;
    }
    case K_CLEAR: f[0] = {1, o.w[2], false, true, true}; return 1;
    case K_CONTRIB: {
        int c0, c1, cd;
        contribution_shape(op, &c0, &c1, &cd);
        const bool first = flags & F_FIRST, init = flags & F_INIT;
        f[0] = {2, (uint32_t)c0, true, false, false};
        f[1] = {3, (uint32_t)c1, true, false, false};
        f[2] = {4, (uint32_t)cd, first && !init, (flags & F_LAST) && !(flags & F_HOLD), first && init};
        return 3;
    }
    case K_PROP:
    case K_IDENTITY: {
        const uint32_t len = kind == K_PROP ? (op == 9 ? 4u : 2u) : o.w[2];
        f[0] = {1, len, !(flags & F_FROMACC), !(flags & F_HOLD), false};
        return 1;
    }
    case K_CLOSURE:
        f[0] = {1, o.w[6], true, false, false};
        f[1] = {2, o.w[6], true, false, false};
        return 2;
    default: return -1;
    }
}

// the current fields of a symjit row (a binding: kernel, inputs, output pairs, direct, the input
// codes, the output components; the codes' layout is the reference's, runner/recurrence_runner.cu:
// 582, 631-684): each parent a code reads is read whole (len 0: the object at the field, whatever
// its size), the destination is created when the row initialises it and its outputs are the
// components 0..m-1 (and no code reads it), else read and written whole; forms the codes read are
// reported in `forms` (their 16-bit fields cannot name LDS). -1 when the binding is out of the table.
constexpr int SYMJIT_FORMS_MAX = 64;  // the momentum inputs one binding may read (`forms` holds them)
inline int symjit_fields(const Instr &o, const std::vector<uint32_t> &bindings, Field *f, uint32_t *forms, int *nforms) {
    const size_t at = o.w[6];
    if (at + 4 > bindings.size()) return -1;
    const uint32_t inputs = bindings[at + 1], pairs = bindings[at + 2];
    if (at + 4 + inputs + pairs > bindings.size() || inputs > (uint32_t)SYMJIT_FORMS_MAX) return -1;
    bool parent[2] = {false, false}, dest_read = false;
    *nforms = 0;
    for (uint32_t k = 0; k < inputs; k++) {
        const uint32_t code = bindings[at + 4 + k], which = (code >> 3) & 1, index = code >> 5;
        switch (code & 7) {
        case SB_PARENT: parent[which] = true; break;
        case SB_DESTINATION: dest_read = true; break;
        case SB_MOMENTUM:
            forms[(*nforms)++] = ((which ? o.w[4] >> 16 : o.w[4] & 0xffff) * 4 + index) / 4;
            break;
        default: break;
        }
    }
    uint32_t m = 0, mask = 0;
    bool distinct = true;
    for (uint32_t j = 0; j < pairs; j++) {
        const uint32_t c = bindings[at + 4 + inputs + j];
        if (c >= 32 || (mask >> c) & 1) distinct = false;
        else mask |= 1u << c;
        m = std::max(m, c + 1);
    }
    const bool create = ((o.w[0] >> 16) & F_INIT) && !dest_read && distinct && pairs > 0 && mask == (1ull << m) - 1;
    int nf = 0;
    if (parent[0]) f[nf++] = {1, 0, true, false, false};
    if (parent[1]) f[nf++] = {2, 0, true, false, false};
    f[nf++] = create ? Field{3, m, false, true, true} : Field{3, 0, true, true, false};
    return nf;
}

// the form fields of row o (word indices)
inline int form_words(const Instr &o, int *w) {
    const int kind = o.w[0] & 0xff, op = (o.w[0] >> 8) & 0xff;
    if (kind == K_CONTRIB && op == 6) { w[0] = 5; w[1] = 6; return 2; }
    if (kind == K_PROP || kind == K_SOURCE) { w[0] = 2; return 1; }
    return 0;
}

struct Obj {
    uint32_t size = 0;      // planes (a form: 2)
    uint32_t lo = 0, hi = 0;  // phases it lives in
    uint64_t uses = 0;      // plane accesses per point
    int room = -1;          // LDS plane, or -1
    uint32_t slot = 0;      // slot plane (currents), else its plan index
    int kind = 0;           // 0 current, 1 amplitude, 2 form
};

// first-fit placement of [lo, hi] x size in occupancy occ[phase][plane] (capacity occ[0].size());
// an object that lives into the reduction (phase P + 1 = occ.size() - 1) also holds its room in
// phase 0: the kernel computes the next unit's forms while it reduces this unit's amplitudes
inline int first_fit(std::vector<std::vector<uint8_t>> &occ, const Obj &o, bool grow) {
    const bool wrap = o.hi + 1 == occ.size();
    for (uint32_t s = 0;; s++) {
        if (s + o.size > occ[0].size()) {
            if (!grow) return -1;
            for (auto &v : occ) v.resize(s + o.size, 0);
        }
        bool free = true;
        for (uint32_t p = o.lo; p <= o.hi && free; p++)
            for (uint32_t q = s; q < s + o.size && free; q++)
                if (occ[p][q]) free = false;
        for (uint32_t q = s; wrap && q < s + o.size && free; q++)
            if (occ[0][q]) free = false;
        if (free) {
            for (uint32_t p = o.lo; p <= o.hi; p++)
                for (uint32_t q = s; q < s + o.size; q++) occ[p][q] = 1;
            for (uint32_t q = s; wrap && q < s + o.size; q++) occ[0][q] = 1;
            return (int)s;
        }
    }
}

}  // namespace coop_detail

// The cooperative program of `ops` (a fused program with forwarding and F_ZERO closures, over
// `planes` current planes, `bindings` its symjit rows' binding table) for W = `waves` wavefronts
// and `lds_cap` LDS planes of `plane_bytes`
// (64 current values: 1024 for 16-byte currents, 512 for the f32 ones; a form takes 2 KB of LDS,
// an amplitude plane 1 KB); `used_amps` marks the amplitude planes the reduction reads. ok is
// false when a row or an access is not one the pass models (accesses to parts of a current), an
// amplitude the reduction reads has no closure, or one's first closure does not start from zero.
// `late`: each task but the sources goes to the last phase before its first reader instead of the
// first after its last producer (the same number of phases; values read far from where they are
// made live shorter). `rank`: the LDS ranking's lifetime weight (0: square root, 1: linear, 2: none).
inline CoopProgram coop_program(const std::vector<Instr> &ops, uint32_t planes, uint32_t amplitudes, uint32_t forms,
                                const std::vector<uint8_t> &used_amps, uint32_t waves, uint32_t lds_cap,
                                uint32_t plane_bytes, const std::vector<uint32_t> &bindings, bool late = false,
                                int rank = 0) {
    using namespace coop_detail;
    CoopProgram cp;
    cp.waves = waves;
    const size_t N = ops.size();
    if (waves == 0 || N == 0 || forms >= FORM_LDS || (plane_bytes != 1024 && plane_bytes != 512)) return cp;
    const uint32_t amp_planes = 1024 / plane_bytes, form_planes = 2048 / plane_bytes;
    // tasks
    std::vector<uint32_t> task_of(N), task_first, task_rows;
    bool open_run = false, hold = false;
    for (size_t n = 0; n < N; n++) {
        const int kind = ops[n].w[0] & 0xff, flags = ops[n].w[0] >> 16;
        const bool extend = (flags & F_FROMACC) || (kind == K_CONTRIB && !(flags & F_FIRST));
        if (extend) {
            if (task_first.empty()) return cp;
            if ((flags & F_FROMACC) && !hold) return cp;
            if (kind == K_CONTRIB && !(flags & F_FIRST) && (!open_run || ops[n].w[4] != ops[n - 1].w[4])) return cp;
            task_rows.back()++;
        } else {
            if (open_run || hold) return cp;
            task_first.push_back((uint32_t)n);
            task_rows.push_back(1);
        }
        task_of[n] = (uint32_t)task_first.size() - 1;
        if (kind == K_CONTRIB) open_run = !(flags & F_LAST);
        hold = flags & F_HOLD;
    }
    if (open_run || hold) return cp;
    const uint32_t T = (uint32_t)task_first.size();
    // objects: currents (by creating write), then amplitudes, then forms
    std::vector<Obj> objs;
    std::vector<int> owner(planes, -1);
    std::vector<std::array<int, 8>> fobj(N);  // object of each current field word
    std::vector<std::array<uint32_t, 8>> moved(N);  // current components each field word reads and writes
    std::vector<int> amp_obj(amplitudes, -1);
    // dependences
    std::vector<std::vector<uint32_t>> dep(T);
    std::vector<int> last_writer;
    std::vector<std::vector<uint32_t>> readers;
    std::vector<int> last_closure(amplitudes, -1);
    auto add_dep = [&](uint32_t t, int d) {
        if (d >= 0 && (uint32_t)d != t) dep[t].push_back((uint32_t)d);
    };
    std::vector<std::vector<uint32_t>> obj_tasks;  // tasks that access each object
    std::vector<uint8_t> slot_forms(forms, 0);     // forms a symjit row reads: they stay in the slot
    for (size_t n = 0; n < N; n++) {
        const Instr &o = ops[n];
        const uint32_t t = task_of[n];
        const int kind = o.w[0] & 0xff;
        fobj[n].fill(-1);
        Field f[3];
        uint32_t sforms[SYMJIT_FORMS_MAX];
        int nsforms = 0;
        const int nf = kind == K_SYMJIT ? symjit_fields(o, bindings, f, sforms, &nsforms) : current_fields(o, f);
        if (nf < 0) return cp;
        for (int k = 0; k < nsforms; k++) {
            if (sforms[k] >= forms) return cp;
            slot_forms[sforms[k]] = 1;
        }
        for (int k = 0; k < nf; k++) {
            const uint32_t base = o.w[f[k].word];
            if (!f[k].create && f[k].len == 0) {  // a symjit row's whole object at the field
                if (base >= planes || owner[base] < 0) return cp;
                f[k].len = objs[owner[base]].size;
            }
            const uint32_t len = f[k].len;
            if (len == 0 || base >= planes || len > planes - base) return cp;
            int id;
            if (f[k].create) {
                id = (int)objs.size();
                Obj ob;
                ob.size = len;
                objs.push_back(ob);
                last_writer.push_back(-1);
                readers.emplace_back();
                obj_tasks.emplace_back();
                for (uint32_t p = base; p < base + len; p++) owner[p] = id;
            } else {
                id = owner[base];
                if (id < 0 || objs[id].size != len || (base > 0 && owner[base - 1] == id)) return cp;
                for (uint32_t p = base; p < base + len; p++)
                    if (owner[p] != id) return cp;  // (part of it was overwritten)
            }
            fobj[n][f[k].word] = id;
        }
        // reads before writes within the row
        for (int k = 0; k < nf; k++) {
            const int id = fobj[n][f[k].word];
            if (f[k].read) {
                add_dep(t, last_writer[id]);
                objs[id].uses += f[k].len;
            }
        }
        moved[n].fill(0);
        for (int k = 0; k < nf; k++)
            moved[n][f[k].word] = (f[k].read ? f[k].len : 0) + (f[k].write ? f[k].len : 0);
        for (int k = 0; k < nf; k++) {
            const int id = fobj[n][f[k].word];
            if (f[k].read) readers[id].push_back(t);
            if (f[k].write || f[k].create) {
                add_dep(t, last_writer[id]);
                for (uint32_t r : readers[id]) add_dep(t, (int)r);
                readers[id].clear();
                last_writer[id] = (int)t;
                if (f[k].write) objs[id].uses += f[k].len;
            }
            obj_tasks[id].push_back(t);
        }
        if (kind == K_CLOSURE) {
            const uint32_t q = o.w[3];
            if (q >= amplitudes) return cp;
            // (an amplitude starts from zero in its first closure: the unit's LDS and slot are
            // not cleared)
            if (last_closure[q] < 0 && !((o.w[0] >> 16) & F_ZERO)) return cp;
            add_dep(t, last_closure[q]);
            last_closure[q] = (int)t;
        }
    }
    const uint32_t current_objs = (uint32_t)objs.size();
    // phases (task phases 1..P; 0 is the kernel's forms phase, P + 1 its reduction)
    std::vector<uint32_t> phase(T, 1);
    uint32_t P = 1;
    for (uint32_t t = 0; t < T; t++) {
        for (uint32_t d : dep[t]) phase[t] = std::max(phase[t], phase[d] + 1);
        P = std::max(P, phase[t]);
    }
    if (late) {  // as late as the readers allow (a task with no reader: the last phase), sources first
        std::vector<uint32_t> latest(T, P);
        for (uint32_t t = T; t-- > 0;)
            for (uint32_t d : dep[t]) latest[d] = std::min(latest[d], latest[t] - 1);
        for (uint32_t t = 0; t < T; t++)
            if ((ops[task_first[t]].w[0] & 0xff) != K_SOURCE) phase[t] = latest[t];
    }
    cp.phases = P;
    // lifetimes of the currents
    for (uint32_t id = 0; id < current_objs; id++) {
        Obj &ob = objs[id];
        ob.lo = P + 1;
        ob.hi = 0;
        for (uint32_t t : obj_tasks[id]) {
            ob.lo = std::min(ob.lo, phase[t]);
            ob.hi = std::max(ob.hi, phase[t]);
        }
    }
    // amplitudes: from their first closure to the reduction
    std::vector<uint32_t> first_closure(amplitudes, UINT32_MAX);
    for (size_t n = 0; n < N; n++)
        if ((ops[n].w[0] & 0xff) == K_CLOSURE) {
            const uint32_t q = ops[n].w[3];
            first_closure[q] = std::min(first_closure[q], phase[task_of[n]]);
        }
    for (uint32_t q = 0; q < amplitudes; q++) {
        const bool used = q < used_amps.size() && used_amps[q];
        if (used && first_closure[q] == UINT32_MAX) return cp;  // (read but never written: would need zeroing)
        if (first_closure[q] == UINT32_MAX) continue;
        Obj ob;
        ob.kind = 1;
        ob.size = amp_planes;
        ob.lo = first_closure[q];
        ob.hi = P + 1;
        ob.slot = q;
        for (size_t n = 0; n < N; n++)
            if ((ops[n].w[0] & 0xff) == K_CLOSURE && ops[n].w[3] == q)
                ob.uses += (((ops[n].w[0] >> 16) & F_ZERO) ? 1 : 2) * amp_planes;
        ob.uses += used ? amp_planes : 0;
        amp_obj[q] = (int)objs.size();
        objs.push_back(ob);
    }
    // forms: from the forms phase to their last reader
    std::vector<int> form_obj(forms, -1);
    for (uint32_t f = 0; f < forms; f++) {
        Obj ob;
        ob.kind = 2;
        ob.size = form_planes;
        ob.lo = 0;
        ob.hi = 0;
        ob.slot = f;
        form_obj[f] = (int)objs.size();
        objs.push_back(ob);
    }
    for (size_t n = 0; n < N; n++) {
        int w[2];
        const int k = form_words(ops[n], w);
        for (int j = 0; j < k; j++) {
            const uint32_t f = ops[n].w[w[j]] & ((1u << FORM_SLOT_SHIFT) - 1);
            if (f >= forms) return cp;
            Obj &ob = objs[form_obj[f]];
            ob.hi = std::max(ob.hi, phase[task_of[n]]);
            ob.uses += form_planes;
        }
    }
    // LDS placement by density, then the slot for the currents left
    std::vector<uint32_t> order(objs.size());
    for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
    auto density = [&](uint32_t i) {
        const Obj &o = objs[i];
        const double life = (double)(o.hi - o.lo + 1);
        return (double)o.uses / o.size / (rank == 0 ? std::sqrt(life) : rank == 1 ? life : 1.0);
    };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return density(a) > density(b); });
    std::vector<std::vector<uint8_t>> lds(P + 2, std::vector<uint8_t>(lds_cap, 0));
    std::vector<std::vector<uint8_t>> slot(P + 2, std::vector<uint8_t>(1, 0));
    for (uint32_t i : order) {
        Obj &o = objs[i];
        if (o.uses == 0) continue;  // (a form no row reads)
        if (o.kind == 2 && slot_forms[o.slot]) continue;
        if (lds_cap) o.room = first_fit(lds, o, false);
        if (o.room >= 0) cp.lds_planes = std::max(cp.lds_planes, (uint32_t)o.room + o.size);
    }
    for (uint32_t i : order) {
        Obj &o = objs[i];
        if (o.kind != 0 || o.room >= 0) continue;
        o.slot = (uint32_t)first_fit(slot, o, true);
        cp.slot_planes = std::max(cp.slot_planes, o.slot + o.size);
    }
    for (uint32_t id = 0; id < current_objs; id++)
        (objs[id].room >= 0 ? cp.lds_accesses : cp.slot_accesses) += objs[id].uses;
    // wavefronts: longest task first to the least-loaded wavefront, per phase; the fetch of the
    // next unit goes to the least-loaded wavefront of the phase where that one's slack (the phase's
    // largest load minus its own) is largest
    // (a task's cost: its arithmetic, plus a charge per current component it moves through the
    // slot or LDS, the slot's being its wait on L2 or beyond)
    std::vector<uint32_t> cost(T, 0), wave_of(T, 0);
    for (size_t n = 0; n < N; n++) {
        cost[task_of[n]] += row_cost(ops[n]);
        for (int w = 1; w < 8; w++)
            if (fobj[n][w] >= 0) cost[task_of[n]] += moved[n][w] * (objs[fobj[n][w]].room >= 0 ? COST_LDS : COST_SLOT);
    }
    uint32_t fetch_phase = 1, fetch_wave = 0;
    uint64_t best_slack = 0;
    std::vector<std::vector<uint64_t>> loads(P + 1);
    for (uint32_t p = 1; p <= P; p++) {
        std::vector<uint32_t> ts;
        for (uint32_t t = 0; t < T; t++)
            if (phase[t] == p) ts.push_back(t);
        std::stable_sort(ts.begin(), ts.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
        std::vector<uint64_t> load(waves, 0);
        for (uint32_t t : ts) {
            const uint32_t w = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
            load[w] += cost[t];
            wave_of[t] = w;
        }
        loads[p] = load;
        const uint32_t least = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
        const uint64_t slack = *std::max_element(load.begin(), load.end()) - load[least];
        if (p == 1 || slack > best_slack) {
            best_slack = slack;
            fetch_phase = p;
            fetch_wave = least;
        }
    }
    // the rewritten rows: device fields, and the oracle's
    auto current_field = [&](int id, bool oracle) {
        const Obj &o = objs[id];
        if (o.room >= 0) return LDS_PLANE + (uint32_t)o.room;
        (void)oracle;
        return o.slot;
    };
    cp.form_field.assign(forms, 0);
    for (uint32_t f = 0; f < forms; f++) {
        const Obj &o = objs[form_obj[f]];
        cp.form_field[f] = o.room >= 0 ? FORM_LDS | (uint32_t)o.room : f;
    }
    cp.amp_field.assign(amplitudes, 0);
    for (uint32_t q = 0; q < amplitudes; q++) {
        const int id = amp_obj[q];
        cp.amp_field[q] = id >= 0 && objs[id].room >= 0 ? LDS_PLANE + (uint32_t)objs[id].room : q;
    }
    std::vector<Instr> dev(N), ora(N);
    for (size_t n = 0; n < N; n++) {
        Instr d = ops[n], r = ops[n];
        for (int w = 1; w < 8; w++)
            if (fobj[n][w] >= 0) {
                d.w[w] = current_field(fobj[n][w], false);
                r.w[w] = d.w[w];
            }
        int fw[2];
        const int k = form_words(ops[n], fw);
        for (int j = 0; j < k; j++) {
            const uint32_t f = ops[n].w[fw[j]] & ((1u << FORM_SLOT_SHIFT) - 1);
            d.w[fw[j]] = cp.form_field[f];
            r.w[fw[j]] = f;
        }
        if ((ops[n].w[0] & 0xff) == K_CLOSURE) d.w[3] = cp.amp_field[ops[n].w[3]];
        dev[n] = d;
        ora[n] = r;
    }
    // the streams
    Instr barrier, fetch;
    memset(&barrier, 0, sizeof barrier);
    barrier.w[0] = K_BARRIER;
    memset(&fetch, 0, sizeof fetch);
    fetch.w[0] = K_FETCH;
    cp.start.assign(waves + 1, 0);
    for (uint32_t w = 0; w < waves; w++) {
        cp.start[w] = (uint32_t)cp.ops.size();
        for (uint32_t p = 1; p <= P; p++) {
            if (p == fetch_phase && w == fetch_wave) cp.ops.push_back(fetch);
            for (uint32_t t = 0; t < T; t++)
                if (phase[t] == p && wave_of[t] == w)
                    for (uint32_t r = 0; r < task_rows[t]; r++) cp.ops.push_back(dev[task_first[t] + r]);
            if (p < P) {
                // (the barrier row also sets the wavefront's priority for the next phase: the most
                // loaded wavefronts, those within a tenth of the largest load, issue first)
                Instr b = barrier;
                const uint64_t top = *std::max_element(loads[p + 1].begin(), loads[p + 1].end());
                b.w[1] = loads[p + 1][w] * 10 >= top * 9 ? 1u : 0u;
                cp.ops.push_back(b);
            }
        }
    }
    cp.start[waves] = (uint32_t)cp.ops.size();
    for (uint32_t p = 1; p <= P; p++)
        for (uint32_t w = 0; w < waves; w++)
            for (uint32_t t = 0; t < T; t++)
                if (phase[t] == p && wave_of[t] == w)
                    for (uint32_t r = 0; r < task_rows[t]; r++) cp.serial.push_back(ora[task_first[t] + r]);
    // checks: every wavefront meets P - 1 barriers, and the streams hold one fetch; the placement
    // keeps one object per room and phase (an object never shares a room with one live in any of
    // its phases)
    uint32_t fetches = 0;
    for (uint32_t w = 0; w < waves; w++) {
        uint32_t b = 0;
        for (uint32_t n = cp.start[w]; n < cp.start[w + 1]; n++) {
            b += (cp.ops[n].w[0] & 0xff) == K_BARRIER;
            fetches += (cp.ops[n].w[0] & 0xff) == K_FETCH;
        }
        if (b != P - 1) return cp;
    }
    if (fetches != 1) return cp;
    // the first phase holds only sources and clears (every other row reads a current), so a kernel
    // may run it with other constants than the rest; it ends at each stream's first barrier
    if (P < 2) return cp;
    cp.split.assign(waves, 0);
    for (uint32_t w = 0; w < waves; w++) {
        uint32_t n = cp.start[w];
        while ((cp.ops[n].w[0] & 0xff) != K_BARRIER) {
            const int kind = cp.ops[n].w[0] & 0xff;
            if (kind != K_SOURCE && kind != K_CLEAR && kind != K_FETCH) return cp;
            n++;
        }
        cp.split[w] = n;
    }
    // (an amplitude's room is also held in phase 0, where the next unit's forms are written)
    std::vector<std::vector<int>> held(P + 2, std::vector<int>(std::max(cp.lds_planes, 1u), -1));
    std::vector<std::vector<int>> held_slot(P + 2, std::vector<int>(std::max(cp.slot_planes, 1u), -1));
    for (uint32_t i = 0; i < objs.size(); i++) {
        const Obj &o = objs[i];
        if (o.uses == 0 || (o.room < 0 && o.kind != 0)) continue;
        auto &h = o.room >= 0 ? held : held_slot;
        const uint32_t s = o.room >= 0 ? (uint32_t)o.room : o.slot;
        for (uint32_t p = o.lo; p <= o.hi; p++)
            for (uint32_t q = s; q < s + o.size; q++) {
                if (q >= h[p].size() || h[p][q] >= 0) return cp;
                h[p][q] = (int)i;
            }
        if (o.hi == P + 1 && o.room >= 0)
            for (uint32_t q = s; q < s + o.size; q++) {
                if (h[0][q] >= 0) return cp;
                h[0][q] = (int)i;
            }
    }
    if (cp.lds_planes > lds_cap) return cp;
    cp.ok = true;
    return cp;
}

}  // namespace pacuda
