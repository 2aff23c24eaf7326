// SPDX-License-Identifier: 0BSD
//
// Host side: a plan's 64-bit words -> the instruction stream the interpreter runs.
//
// Three stages, each a function of its own:
//   decode_rows     the words into rows, by a table of field layouts per op (the word layout of
//                   runner/recurrence_runner.cu:10-36 and convert/src/program.rs:88-108: op in
//                   bits 0-3, then the op's fields LSB first with the plan's widths W_P..W_K or
//                   fixed widths); a binding (op 14) carries its codes in the data words after it
//   sort_runs       each maximal run of contribution rows (ops 2-6) stably by destination, when
//                   the destinations are disjoint and no row reads a destination of the run
//                   (the rule of runner/recurrence_runner.cu:977-1007: every destination keeps
//                   its own order of rows, so its sum keeps its bits)
//   emit_program    one 32-byte instruction per row (bindings go to a table of their own), the
//                   first and last rows of each destination marked (F_FIRST / F_LAST, so a lane
//                   keeps the partial sum in registers), every table extent the program touches
//                   recorded, symjit rows checked against the kernel table
// compile_program runs the three and reports the first error as a message.
#pragma once
#include <stdint.h>

#include <vector>

#include "pacman_recurrence.h"
#include "program.hpp"

namespace pacuda {

// one decoded row: its op and fields in word order; a binding's codes are data[data_at ..
// data_at + data_len) of its RowSet
struct PlanRow {
    uint8_t op = 0;
    uint32_t f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t data_at = 0, data_len = 0;
};

struct RowSet {
    std::vector<PlanRow> rows;
    std::vector<uint32_t> data;
    const char *error = nullptr;
};

RowSet decode_rows(const uint64_t *words, uint32_t count, const uint8_t *width);
void sort_runs(std::vector<PlanRow> &rows);

// The generated symjit kernels a runner carries: how many, and the largest input and output
// counts (doubles) one may take.
struct KernelLimits {
    uint32_t count, max_in, max_out;
};

// The program of a plan. `bindings` holds each symjit binding as the interpreter reads it:
// kernel, inputs, output pairs, direct, the input codes, the output components; a K_SYMJIT
// instruction names its binding by offset. need[t]: one more than the largest index of table t
// the program touches.
struct CompiledProgram {
    std::vector<Instr> ops;
    std::vector<uint32_t> bindings;
    uint32_t sources = 0;
    uint64_t need[N_COUNT] = {0, 0, 0, 0, 0};
    bool symjit = false;
    const char *error = nullptr;
};

CompiledProgram compile_program(const uint64_t *words, uint32_t count, const uint8_t *width, bool sort,
                                const KernelLimits &kernels);

// components of parent 0, parent 1 and the destination of a contribution op (2-6)
struct ContributionShape {
    int c0, c1, cd;
};
ContributionShape contribution_shape(int op);
inline void contribution_shape(int op, int *c0, int *c1, int *cd) {
    const ContributionShape s = contribution_shape(op);
    *c0 = s.c0;
    *c1 = s.c1;
    *cd = s.cd;
}

}  // namespace pacuda
