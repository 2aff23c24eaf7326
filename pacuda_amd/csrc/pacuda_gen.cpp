// SPDX-License-Identifier: 0BSD
//
// pacuda-gen: HIP/C++ functions for the runner's symjit rows (op 13) from plane-kernel files.
//
//   pacuda-gen hip FILE.pk... [-o OUT.hpp]
//
// The runner's contract with generated kernels (the one runner/recurrence_runner.cu:554-573 and
// gen/src/main.rs:47-99 set up): one function per distinct kernel file (a file whose bytes
// repeat an earlier one's is skipped), `pacuda_kernel(k, simd, in, out)` dispatching on the
// kernel index, PACUDA_KERNEL_COUNT, PACUDA_MAX_IN (at least 1), PACUDA_MAX_OUT (at least 2),
// PACUDA_SIMD_LANES, and `pacuda_kernel_sha256[k]`, the lower-case hex SHA-256 of kernel k's
// file, by which a recording names its kernels (pacman_recurrence_v1_symjit_kernel).
//
// A plane-kernel file is straight-line f64 code, one operation per line:
//   in N                   the kernel reads in[0 .. N)
//   out M                  and writes out[0 .. M) (pairs: re, im)
//   NAME = OP ARGS...      OP: add a b, sub a b, mul a b, div a b (a / b), fma a b c (a * b + c,
//                          one rounding), neg a, sqrt a, copy a, bits HEX16 (a constant by its
//                          64-bit pattern)
//   # ...                  a comment
// Arguments are in<i> or names defined above; `out<j> = ...` sets an output (each once). Every
// operation is one rounded IEEE operation in the emitted code (built with FP contraction off), so
// the functions give the same bits on the CPU and on gfx950. symjit's own plane files (MIR) are
// not read here: this image has neither symjit nor files of it.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

namespace {

// ---- SHA-256 (FIPS 180-4) ------------------------------------------------------------------

uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

std::string sha256_hex(const std::string &bytes) {
    static const uint32_t K[64] = {
        0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
        0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
        0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
        0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
        0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
        0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
        0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
        0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
    uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    std::string m = bytes;
    const uint64_t bit_length = (uint64_t)bytes.size() * 8;
    m.push_back((char)0x80);
    while (m.size() % 64 != 56) m.push_back(0);
    for (int k = 7; k >= 0; k--) m.push_back((char)(bit_length >> (8 * k)));
    for (size_t block = 0; block < m.size(); block += 64) {
        uint32_t w[64];
        for (int t = 0; t < 16; t++)
            w[t] = (uint32_t)(uint8_t)m[block + 4 * t] << 24 | (uint32_t)(uint8_t)m[block + 4 * t + 1] << 16 |
                   (uint32_t)(uint8_t)m[block + 4 * t + 2] << 8 | (uint32_t)(uint8_t)m[block + 4 * t + 3];
        for (int t = 16; t < 64; t++) {
            const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
            const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
            w[t] = w[t - 16] + s0 + w[t - 7] + s1;
        }
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
        for (int t = 0; t < 64; t++) {
            const uint32_t t1 = hh + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + K[t] + w[t];
            const uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
            hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
        }
        h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
    }
    char out[65];
    for (int k = 0; k < 8; k++) snprintf(out + 8 * k, 9, "%08x", h[k]);
    return std::string(out, 64);
}

// ---- plane-kernel files --------------------------------------------------------------------

struct Kernel {
    std::string name, digest, body;
    unsigned in = 0, out = 0;
};

[[noreturn]] void fail(const std::string &where, const std::string &what) {
    fprintf(stderr, "pacuda-gen: %s: %s\n", where.c_str(), what.c_str());
    exit(1);
}

std::string function_name(const std::string &path) {
    std::string stem = path.substr(path.find_last_of('/') + 1);
    if (stem.size() > 3 && stem.compare(stem.size() - 3, 3, ".pk") == 0) stem.resize(stem.size() - 3);
    for (char &c : stem)
        if (!isalnum((unsigned char)c)) c = '_';
    return "pk_" + stem;
}

Kernel read_kernel(const std::string &path, const std::string &bytes) {
    Kernel k;
    k.name = function_name(path);
    k.digest = sha256_hex(bytes);
    std::istringstream lines(bytes);
    std::string line;
    std::set<std::string> defined;
    std::set<unsigned> written;
    int number = 0;
    bool have_in = false, have_out = false;
    std::ostringstream body;
    auto where = [&]() { return path + ":" + std::to_string(number); };
    auto operand = [&](const std::string &a) -> std::string {
        if (a.size() > 2 && a.compare(0, 2, "in") == 0 && isdigit((unsigned char)a[2])) {
            const unsigned i = (unsigned)std::stoul(a.substr(2));
            if (i >= k.in) fail(where(), "input " + a + " beyond `in`");
            return "in[" + std::to_string(i) + "]";
        }
        if (!defined.count(a)) fail(where(), "undefined name " + a);
        return "v_" + a;
    };
    while (std::getline(lines, line)) {
        number++;
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream t(line);
        std::vector<std::string> w;
        for (std::string s; t >> s;) w.push_back(s);
        if (w.empty()) continue;
        if (w.size() == 2 && w[0] == "in" && !have_in) {
            k.in = (unsigned)std::stoul(w[1]);
            have_in = true;
            continue;
        }
        if (w.size() == 2 && w[0] == "out" && !have_out) {
            k.out = (unsigned)std::stoul(w[1]);
            have_out = true;
            continue;
        }
        if (!have_in || !have_out) fail(where(), "`in` and `out` come first");
        if (w.size() < 3 || w[1] != "=") fail(where(), "expected NAME = OP ARGS");
        const std::string &op = w[2];
        std::vector<std::string> a(w.begin() + 3, w.end());
        static const std::map<std::string, size_t> ARITY = {{"add", 2}, {"sub", 2}, {"mul", 2}, {"div", 2}, {"fma", 3},
                                                            {"neg", 1}, {"sqrt", 1}, {"copy", 1}, {"bits", 1}};
        auto it = ARITY.find(op);
        if (it == ARITY.end()) fail(where(), "unknown operation " + op);
        if (a.size() != it->second) fail(where(), op + " takes " + std::to_string(it->second) + " arguments");
        std::string expr;
        if (op == "bits") {
            if (a[0].size() != 16 || a[0].find_first_not_of("0123456789abcdefABCDEF") != std::string::npos)
                fail(where(), "bits takes 16 hex digits");
            expr = "pk_bits(0x" + a[0] + "ull)";
        } else if (op == "add") expr = operand(a[0]) + " + " + operand(a[1]);
        else if (op == "sub") expr = operand(a[0]) + " - " + operand(a[1]);
        else if (op == "mul") expr = operand(a[0]) + " * " + operand(a[1]);
        else if (op == "div") expr = operand(a[0]) + " / " + operand(a[1]);
        else if (op == "fma") expr = "__builtin_fma(" + operand(a[0]) + ", " + operand(a[1]) + ", " + operand(a[2]) + ")";
        else if (op == "neg") expr = "-" + operand(a[0]);
        else if (op == "sqrt") expr = "__builtin_sqrt(" + operand(a[0]) + ")";
        else expr = operand(a[0]);
        const std::string &dst = w[0];
        if (dst.size() > 3 && dst.compare(0, 3, "out") == 0 && isdigit((unsigned char)dst[3])) {
            const unsigned j = (unsigned)std::stoul(dst.substr(3));
            if (j >= k.out) fail(where(), "output " + dst + " beyond `out`");
            if (!written.insert(j).second) fail(where(), dst + " written twice");
            body << "    out[" << j << "] = " << expr << ";\n";
        } else {
            if (dst.compare(0, 2, "in") == 0 || !isalpha((unsigned char)dst[0])) fail(where(), "bad name " + dst);
            if (!defined.insert(dst).second) fail(where(), dst + " defined twice");
            body << "    const double v_" << dst << " = " << expr << ";\n";
        }
    }
    if (!have_in || !have_out) fail(path, "no `in` / `out`");
    if (written.size() != k.out) fail(path, "not every output is written");
    k.body = body.str();
    return k;
}

std::string emit(const std::vector<Kernel> &kernels) {
    std::ostringstream s;
    unsigned max_in = 1, max_out = 2;
    for (const Kernel &k : kernels) {
        max_in = std::max(max_in, k.in);
        max_out = std::max(max_out, k.out);
    }
    s << "// Generated by pacuda-gen from plane-kernel files; do not edit.\n"
         "#pragma once\n"
         "#include <stdint.h>\n\n"
         "#ifndef PACUDA_KFN\n"
         "#if defined(__HIPCC__)\n"
         "#define PACUDA_KFN __host__ __device__ __attribute__((always_inline)) inline\n"
         "#else\n"
         "#define PACUDA_KFN inline\n"
         "#endif\n"
         "#endif\n\n"
         "PACUDA_KFN double pk_bits(unsigned long long u) { return __builtin_bit_cast(double, u); }\n\n"
      << "#define PACUDA_KERNEL_COUNT " << kernels.size() << "\n"
      << "#define PACUDA_MAX_IN " << max_in << "\n"
      << "#define PACUDA_MAX_OUT " << max_out << "\n"
      << "#define PACUDA_SIMD_LANES 0\n"
      << "static const char *const pacuda_kernel_sha256[" << std::max<size_t>(kernels.size(), 1) << "] = {\n";
    for (const Kernel &k : kernels) s << "    \"" << k.digest << "\",  // " << k.name << "\n";
    if (kernels.empty()) s << "    \"\",\n";
    s << "};\n";
    for (const Kernel &k : kernels)
        s << "\n// " << k.in << " inputs, " << k.out << " outputs\n"
          << "PACUDA_KFN void " << k.name << "(const double *in, double *out) {\n"
          << k.body << "}\n";
    s << "\nPACUDA_KFN void pacuda_kernel(uint32_t k, bool simd, const double *in, double *out) {\n"
         "    (void)simd;\n"
         "    switch (k) {\n";
    for (size_t n = 0; n < kernels.size(); n++) s << "    case " << n << ": " << kernels[n].name << "(in, out); break;\n";
    s << "    default: (void)in; (void)out; break;\n"
         "    }\n"
         "}\n";
    return s.str();
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2 || strcmp(argv[1], "hip") != 0) {
        fprintf(stderr, "usage: pacuda-gen hip FILE.pk... [-o OUT.hpp]\n");
        return 2;
    }
    std::vector<std::string> paths;
    std::string output;
    for (int a = 2; a < argc; a++) {
        if (!strcmp(argv[a], "-o") && a + 1 < argc) output = argv[++a];
        else paths.push_back(argv[a]);
    }
    std::vector<Kernel> kernels;
    std::set<std::string> digests, names;
    for (const std::string &p : paths) {
        std::ifstream f(p, std::ios::binary);
        if (!f) fail(p, "cannot read");
        std::ostringstream b;
        b << f.rdbuf();
        Kernel k = read_kernel(p, b.str());
        if (!digests.insert(k.digest).second) continue;  // the same kernel again
        if (!names.insert(k.name).second) fail(p, "another kernel is named " + k.name);
        kernels.push_back(k);
    }
    const std::string source = emit(kernels);
    if (output.empty()) {
        fputs(source.c_str(), stdout);
    } else {
        std::ofstream o(output, std::ios::binary);
        o << source;
        if (!o) fail(output, "cannot write");
    }
    return 0;
}
