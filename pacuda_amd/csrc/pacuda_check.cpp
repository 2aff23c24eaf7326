// SPDX-License-Identifier: 0BSD
//
// pacuda-check FILE.pacuda --runner LIBRARY [--kernel f64|df64|mixed|mixed64] [--tile N]
//              [--host-reduce] [--key] [--dump] [--examples EXAMPLES.json]
// pacuda-check FILE.pacuda --write-examples EXAMPLES.json
//
// Reference points come from the recording, or from an examples.json in pac2json's format
// (convert/src/pac2json.rs:774-804, read back as :1137-1190): {"description": ..., "programs":
// [{"process": key, "colour": id, "points": [{"momenta": [[E, px, py, pz] per leg, before the
// crossing], "outputs": [per helicity, group factor applied], "total": their sum}]}]}, one
// program per (process, colour) in the recording's order; --write-examples writes the
// recording's points in that format (numbers in their shortest round-trip form).
//
// Replays the reference points of every (process, colour) of a .pacuda v2 recording through a
// runner library's C ABI and compares them with the recorded CPU values, as
// `pacuda-convert check` does (convert/src/main.rs:100-126, convert/src/replay.rs:156-298):
// the momenta crossed into source-major tiles, a FORMS_ON_DEVICE run (KEEP unless
// --host-reduce), the resolved reduction into [point][helicity] (or the host's), the group
// factor; then the count of values differing in bits, the worst error relative to the point's
// total and the worst relative error of a point's total. With the f64 kernel any differing
// value is a failure (exit 1). replay.rs's calls carry no key; with --key each colour's calls
// carry one (the FNV-1a hash of its words and widths), as a caller that keeps its programs does. The file grammar is convert/src/format.rs:6-29 (little endian,
// no padding, counted arrays, u32-length strings); the reader rejects trailing bytes and the
// inconsistencies umami/umami_pacuda.c:253-257 rejects.
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <charconv>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pacman_recurrence.h"

namespace {

struct Reader {
    const std::vector<uint8_t> &b;
    size_t at = 0;
    explicit Reader(const std::vector<uint8_t> &bytes) : b(bytes) {}
    const uint8_t *take(size_t n) {
        if (n > b.size() - at) throw std::runtime_error("truncated file");
        const uint8_t *p = b.data() + at;
        at += n;
        return p;
    }
    template <class T> T get() {
        T v;
        memcpy(&v, take(sizeof v), sizeof v);
        return v;
    }
    uint32_t u32() { return get<uint32_t>(); }
    std::string str() {
        uint32_t n = u32();
        const uint8_t *p = take(n);
        return std::string((const char *)p, n);
    }
    template <class T> std::vector<T> vec(size_t n) {
        if (n > (b.size() - at) / sizeof(T)) throw std::runtime_error("truncated file");
        std::vector<T> v(n);
        if (n) memcpy(v.data(), take(n * sizeof(T)), n * sizeof(T));
        return v;
    }
    std::vector<uint32_t> runs() { return vec<uint32_t>(2 * (size_t)u32()); }
};

struct Colour {
    std::string id;
    uint32_t tile;
    std::vector<uint64_t> words;
    uint8_t width[8];
    uint32_t current_planes, amplitude_planes, momentum_planes;
    std::vector<double> scale_re, scale_im, factor_re, factor_im;
    std::vector<uint32_t> zero_runs, download_runs;
    uint32_t external_count;
    std::vector<uint32_t> form_terms, term_slots;
    std::vector<double> term_coefficients;
    std::vector<uint32_t> offsets, planes;
    std::vector<double> weights;
    uint32_t points;
    std::vector<double> momenta, values;
};

struct Process {
    std::string key;
    uint32_t legs, helicities;
    std::string alpha_s_name;
    double alpha_s;
    int32_t qcd_power;
    int64_t num, den;
    std::vector<uint32_t> cross_leg;
    std::vector<double> cross_sign;
    std::vector<Colour> colours;
};

struct File {
    uint32_t kernel, batch;
    std::vector<std::string> sha256;
    std::vector<Process> processes;
};

File read_file(const std::vector<uint8_t> &bytes) {
    Reader r(bytes);
    static const uint8_t magic[8] = {'P', 'A', 'C', 'U', 'D', 'A', 0, 1};
    if (memcmp(r.take(8), magic, 8)) throw std::runtime_error("not a .pacuda file");
    if (r.u32() != 2) throw std::runtime_error("not a version 2 .pacuda file");
    File f;
    f.kernel = r.u32();
    f.batch = r.u32();
    for (uint32_t n = r.u32(), k = 0; k < n; k++) f.sha256.emplace_back((const char *)r.take(64), 64);
    for (uint32_t np = r.u32(), k = 0; k < np; k++) {
        Process p;
        p.key = r.str();
        p.legs = r.u32();
        p.helicities = r.u32();
        p.alpha_s_name = r.str();
        p.alpha_s = r.get<double>();
        p.qcd_power = r.get<int32_t>();
        for (uint32_t n = r.u32(), j = 0; j < n; j++) {
            r.str();
            r.get<double>();
            r.get<double>();
        }
        p.num = r.get<int64_t>();
        p.den = r.get<int64_t>();
        uint32_t crossing = r.u32();
        if (crossing && crossing != p.legs) throw std::runtime_error("process " + p.key + ": a crossing of the wrong length");
        for (uint32_t j = 0; j < crossing; j++) {
            p.cross_leg.push_back(r.u32());
            p.cross_sign.push_back(r.get<double>());
            if (p.cross_leg.back() >= p.legs) throw std::runtime_error("process " + p.key + ": a crossing leg out of range");
        }
        for (uint32_t nc = r.u32(), j = 0; j < nc; j++) {
            Colour c;
            c.id = r.str();
            c.tile = r.u32();
            c.words = r.vec<uint64_t>(r.u32());
            memcpy(c.width, r.take(8), 8);
            c.current_planes = r.u32();
            c.amplitude_planes = r.u32();
            c.momentum_planes = r.u32();
            uint32_t scales = r.u32();
            c.scale_re = r.vec<double>(scales);
            c.scale_im = r.vec<double>(scales);
            uint32_t factors = r.u32();
            c.factor_re = r.vec<double>(factors);
            c.factor_im = r.vec<double>(factors);
            c.zero_runs = r.runs();
            c.download_runs = r.runs();
            c.external_count = r.u32();
            c.form_terms = r.runs();
            uint32_t terms = r.u32();
            c.term_slots = r.vec<uint32_t>(terms);
            c.term_coefficients = r.vec<double>(terms);
            uint32_t outputs = r.u32();
            c.offsets = r.vec<uint32_t>((size_t)outputs + 1);
            c.planes = r.vec<uint32_t>(c.offsets.back());
            c.weights = r.vec<double>(c.offsets.back());
            c.points = r.u32();
            c.momenta = r.vec<double>((size_t)c.points * p.legs * 4);
            c.values = r.vec<double>((size_t)c.points * p.helicities);
            if (outputs != p.helicities || c.external_count != p.legs || c.tile == 0 || c.offsets[0] != 0)
                throw std::runtime_error("process " + p.key + " colour " + c.id + ": inconsistent template");
            for (uint32_t o = 0; o < outputs; o++)
                if (c.offsets[o] > c.offsets[o + 1]) throw std::runtime_error("colour " + c.id + ": offsets not monotone");
            if (c.form_terms.size() * 2 != (size_t)c.momentum_planes) throw std::runtime_error("colour " + c.id + ": forms");
            p.colours.push_back(std::move(c));
        }
        f.processes.push_back(std::move(p));
    }
    if (r.at != bytes.size()) throw std::runtime_error("trailing bytes");
    return f;
}

// ---- examples.json ----------------------------------------------------------------------------

// A JSON value: null, bool, number, string, array or object (keys in file order).
struct Json {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<Json> items;
    std::vector<std::pair<std::string, Json>> fields;
    const Json &at(const std::string &key) const {
        for (const auto &f : fields)
            if (f.first == key) return f.second;
        throw std::runtime_error("examples: no \"" + key + "\"");
    }
    const std::vector<Json> &array() const {
        if (kind != ARR) throw std::runtime_error("examples: an array expected");
        return items;
    }
    double number() const {
        if (kind == NUL) return NAN;  // (pac2json writes non-finite numbers as null)
        if (kind != NUM) throw std::runtime_error("examples: a number expected");
        return num;
    }
    const std::string &text() const {
        if (kind != STR) throw std::runtime_error("examples: a string expected");
        return str;
    }
};

class JsonParser {
  public:
    explicit JsonParser(const std::string &t) : t_(t) {}
    Json document() {
        Json v = value();
        space();
        if (at_ != t_.size()) fail("trailing characters");
        return v;
    }

  private:
    const std::string &t_;
    size_t at_ = 0;
    [[noreturn]] void fail(const char *what) {
        throw std::runtime_error(std::string("examples: ") + what + " at byte " + std::to_string(at_));
    }
    void space() {
        while (at_ < t_.size() && (t_[at_] == ' ' || t_[at_] == '\n' || t_[at_] == '\r' || t_[at_] == '\t')) at_++;
    }
    bool eat(char c) {
        space();
        if (at_ < t_.size() && t_[at_] == c) {
            at_++;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail("unexpected character");
    }
    std::string string() {
        expect('"');
        std::string out;
        while (at_ < t_.size() && t_[at_] != '"') {
            char c = t_[at_++];
            if (c != '\\') {
                out.push_back(c);
                continue;
            }
            if (at_ >= t_.size()) fail("a bad escape");
            c = t_[at_++];
            switch (c) {
            case 'n': out.push_back('\n'); break;
            case 't': out.push_back('\t'); break;
            case 'r': out.push_back('\r'); break;
            case 'b': out.push_back('\b'); break;
            case 'f': out.push_back('\f'); break;
            case 'u': {
                if (at_ + 4 > t_.size()) fail("a bad escape");
                unsigned cp = (unsigned)std::stoul(t_.substr(at_, 4), nullptr, 16);
                at_ += 4;
                // UTF-8 (the descriptions hold a few non-ASCII characters; no surrogate pairs)
                if (cp < 0x80) {
                    out.push_back((char)cp);
                } else if (cp < 0x800) {
                    out.push_back((char)(0xc0 | cp >> 6));
                    out.push_back((char)(0x80 | (cp & 0x3f)));
                } else {
                    out.push_back((char)(0xe0 | cp >> 12));
                    out.push_back((char)(0x80 | ((cp >> 6) & 0x3f)));
                    out.push_back((char)(0x80 | (cp & 0x3f)));
                }
                break;
            }
            default: out.push_back(c); break;
            }
        }
        if (at_ >= t_.size()) fail("an unterminated string");
        at_++;
        return out;
    }
    Json value() {
        space();
        if (at_ >= t_.size()) fail("a value expected");
        Json v;
        const char c = t_[at_];
        if (c == '{') {
            at_++;
            v.kind = Json::OBJ;
            if (eat('}')) return v;
            do {
                std::string key = string();
                expect(':');
                v.fields.emplace_back(key, value());
            } while (eat(','));
            expect('}');
        } else if (c == '[') {
            at_++;
            v.kind = Json::ARR;
            if (eat(']')) return v;
            do v.items.push_back(value());
            while (eat(','));
            expect(']');
        } else if (c == '"') {
            v.kind = Json::STR;
            v.str = string();
        } else if (t_.compare(at_, 4, "null") == 0) {
            at_ += 4;
        } else if (t_.compare(at_, 4, "true") == 0) {
            at_ += 4;
            v.kind = Json::BOOL;
            v.b = true;
        } else if (t_.compare(at_, 5, "false") == 0) {
            at_ += 5;
            v.kind = Json::BOOL;
        } else {
            // strtod rounds correctly: the shortest form a writer prints reads back to its bits
            char *end = nullptr;
            v.kind = Json::NUM;
            v.num = strtod(t_.c_str() + at_, &end);
            if (end == t_.c_str() + at_) fail("a value expected");
            at_ = end - t_.c_str();
        }
        return v;
    }
};

std::string read_text(const char *path) {
    FILE *fp = fopen(path, "rb");
    if (!fp) throw std::runtime_error(std::string(path) + ": cannot open");
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, got);
    fclose(fp);
    return text;
}

// the recording's reference points replaced by those of an examples.json
void take_examples(File &f, const Json &doc) {
    const std::vector<Json> &programs = doc.at("programs").array();
    size_t n = 0;
    for (Process &p : f.processes)
        for (Colour &c : p.colours) {
            if (n >= programs.size()) throw std::runtime_error("examples: fewer programs than the recording");
            const Json &e = programs[n++];
            if (e.at("process").text() != p.key || e.at("colour").text() != c.id)
                throw std::runtime_error("examples: program " + std::to_string(n - 1) + " is not " + p.key + " " + c.id);
            c.momenta.clear();
            c.values.clear();
            c.points = 0;
            for (const Json &point : e.at("points").array()) {
                const std::vector<Json> &legs = point.at("momenta").array();
                const std::vector<Json> &outputs = point.at("outputs").array();
                if (legs.size() != p.legs || outputs.size() != p.helicities)
                    throw std::runtime_error("examples: a point of " + p.key + " " + c.id + " of the wrong size");
                for (const Json &m : legs) {
                    if (m.array().size() != 4) throw std::runtime_error("examples: a momentum of the wrong size");
                    for (const Json &x : m.array()) c.momenta.push_back(x.number());
                }
                for (const Json &x : outputs) c.values.push_back(x.number());
                c.points++;
            }
        }
    if (n != programs.size()) throw std::runtime_error("examples: more programs than the recording");
}

std::string json_number(double x) {
    if (!std::isfinite(x)) return "null";
    char buf[64];
    const auto r = std::to_chars(buf, buf + sizeof buf, x);  // the shortest form that reads back to x
    std::string s(buf, r.ptr);
    if (s.find_first_of(".en") == std::string::npos) s += ".0";  // (a float, as pac2json prints one)
    return s;
}

std::string json_string(const std::string &s) {
    std::string out = "\"";
    for (unsigned char c : s) {
        if (c == '"' || c == '\\') {
            out.push_back('\\');
            out.push_back((char)c);
        } else if (c < 0x20) {
            char buf[8];
            snprintf(buf, sizeof buf, "\\u%04x", c);
            out += buf;
        } else {
            out.push_back((char)c);
        }
    }
    return out + "\"";
}

void write_examples(const File &f, const char *path) {
    std::string out = "{\"description\": " +
                      json_string("per program, phase-space points and the CPU's outputs (f64): momenta[leg] = [E, px, py, "
                                  "pz], the caller's legs before the crossing; outputs[h] per helicity, group factor "
                                  "applied; total = the sum of outputs (in order)") +
                      ",\n \"programs\": [";
    bool first_program = true;
    for (const Process &p : f.processes)
        for (const Colour &c : p.colours) {
            out += first_program ? "\n  " : ",\n  ";
            first_program = false;
            out += "{\"process\": " + json_string(p.key) + ", \"colour\": " + json_string(c.id) + ", \"points\": [";
            for (uint32_t q = 0; q < c.points; q++) {
                out += q ? ",\n    " : "\n    ";
                out += "{\"momenta\": [";
                for (uint32_t leg = 0; leg < p.legs; leg++) {
                    out += leg ? ", [" : "[";
                    for (int k = 0; k < 4; k++)
                        out += (k ? ", " : "") + json_number(c.momenta[((size_t)q * p.legs + leg) * 4 + k]);
                    out += "]";
                }
                out += "], \"outputs\": [";
                double total = 0.0;
                for (uint32_t h = 0; h < p.helicities; h++) {
                    const double v = c.values[(size_t)q * p.helicities + h];
                    out += (h ? ", " : "") + json_number(v);
                    total = total + v;
                }
                out += "], \"total\": " + json_number(total) + "}";
            }
            out += "]}";
        }
    out += "\n]}\n";
    FILE *fp = fopen(path, "wb");
    if (!fp || fwrite(out.data(), 1, out.size(), fp) != out.size()) throw std::runtime_error(std::string(path) + ": cannot write");
    fclose(fp);
}

struct Runner {
    int (*run)(const pacman_recurrence_call *);
    int (*reduce_resolved)(const pacman_recurrence_reduce_resolved *);
    int (*configure)(const pacman_recurrence_options *);
    int (*symjit_kernel)(const char *);
};

// op 14 kernel indices rewritten from the file's table to the runner's (convert/src/replay.rs:133-153)
std::vector<uint64_t> remap_kernels(const std::vector<uint64_t> &words, const std::vector<uint32_t> &map) {
    std::vector<uint64_t> out = words;
    for (size_t n = 0; n < out.size();) {
        uint64_t w = out[n];
        if ((w & 15) == 14) {
            uint64_t kernel = (w >> 4) & 0xffff, codes = ((w >> 20) & 0xff) + ((w >> 28) & 0xff);
            if (kernel >= map.size()) throw std::runtime_error("op 14 names a kernel outside the file's table");
            out[n] = (w & ~(0xffffull << 4)) | (uint64_t)map[kernel] << 4;
            n += 1 + (codes + 1) / 2;
        } else {
            n++;
        }
    }
    return out;
}

// a nonzero key for a program: the 64-bit FNV-1a hash of its words and widths
uint64_t program_key(const std::vector<uint64_t> &words, const uint8_t *width) {
    uint64_t hash = 0xcbf29ce484222325ull;
    auto mix = [&](const void *data, size_t bytes) {
        for (size_t k = 0; k < bytes; k++) hash = (hash ^ ((const uint8_t *)data)[k]) * 0x100000001b3ull;
    };
    mix(words.data(), words.size() * sizeof(uint64_t));
    mix(width, 8);
    return hash ? hash : 1;
}

std::vector<double> evaluate(const Runner &rn, const std::vector<uint32_t> &map, const Process &p, const Colour &c,
                             size_t tile, bool host_reduce, bool keyed) {
    const size_t n = p.legs, h = p.helicities, points = c.points;
    std::vector<uint64_t> words = remap_kernels(c.words, map);
    const uint64_t key = keyed ? program_key(words, c.width) : 0;
    std::vector<double> values(points * h, 0.0);
    const size_t stride = tile;
    std::vector<double> cre(c.current_planes * stride), cim(cre.size()), are(c.amplitude_planes * stride), aim(are.size()),
        mom(std::max<size_t>(1, c.momentum_planes * stride)), ext(n * 4 * tile);
    for (size_t start = 0; start < points; start += tile) {
        const size_t count = std::min(tile, points - start);
        for (size_t q = 0; q < count; q++)
            for (size_t leg = 0; leg < n; leg++) {
                size_t src = p.cross_leg.empty() ? leg : p.cross_leg[leg];
                double sign = p.cross_sign.empty() ? 1.0 : p.cross_sign[leg];
                for (size_t k = 0; k < 4; k++) ext[(leg * 4 + k) * count + q] = sign * c.momenta[((start + q) * n + src) * 4 + k];
            }
        pacman_recurrence_call call;
        memset(&call, 0, sizeof call);
        call.abi_version = PACMAN_RECURRENCE_ABI;
        call.struct_size = sizeof call;
        call.words = words.data();
        call.words_key = key;
        call.word_count = (uint32_t)words.size();
        memcpy(call.width, c.width, 8);
        call.current_re = cre.data();
        call.current_im = cim.data();
        call.amplitude_re = are.data();
        call.amplitude_im = aim.data();
        call.momenta = mom.data();
        call.scale_re = c.scale_re.data();
        call.scale_im = c.scale_im.data();
        call.factor_re = c.factor_re.data();
        call.factor_im = c.factor_im.data();
        call.download_amplitude = c.download_runs.data();
        call.current_planes = c.current_planes;
        call.amplitude_planes = c.amplitude_planes;
        call.momentum_planes = c.momentum_planes;
        call.scale_count = (uint32_t)c.scale_re.size();
        call.factor_count = (uint32_t)c.factor_re.size();
        call.download_amplitude_runs = (uint32_t)(c.download_runs.size() / 2);
        call.stride = (uint32_t)stride;
        call.points = (uint32_t)count;
        call.zero_amplitude = c.zero_runs.data();
        call.zero_amplitude_runs = (uint32_t)(c.zero_runs.size() / 2);
        call.flags = PACMAN_RECURRENCE_FORMS_ON_DEVICE | (host_reduce ? 0 : PACMAN_RECURRENCE_KEEP_AMPLITUDES);
        call.external_momenta = ext.data();
        call.external_count = (uint32_t)n;
        call.term_count = (uint32_t)c.term_slots.size();
        call.form_terms = c.form_terms.data();
        call.term_slots = c.term_slots.data();
        call.term_coefficients = c.term_coefficients.data();
        int st = rn.run(&call);
        if (st != 0) throw std::runtime_error("the runner returned " + std::to_string(st));
        if (host_reduce) {
            for (size_t o = 0; o < h; o++)
                for (uint32_t k = c.offsets[o]; k < c.offsets[o + 1]; k++) {
                    const size_t plane = (size_t)c.planes[k] * stride;
                    for (size_t q = 0; q < count; q++) {
                        double re = are[plane + q], im = aim[plane + q];
                        values[(start + q) * h + o] += c.weights[k] * fma(re, re, im * im);
                    }
                }
        } else {
            pacman_recurrence_reduce_resolved r;
            memset(&r, 0, sizeof r);
            r.abi_version = PACMAN_RECURRENCE_ABI;
            r.struct_size = sizeof r;
            r.planes = c.planes.data();
            r.weights = c.weights.data();
            r.offsets = c.offsets.data();
            r.output = values.data() + start * h;
            r.outputs = (uint32_t)h;
            r.points = (uint32_t)count;
            r.output_stride = 1;
            r.point_stride = (uint32_t)h;
            st = rn.reduce_resolved(&r);
            if (st != 0) throw std::runtime_error("the resolved reduction returned " + std::to_string(st));
        }
    }
    if (p.num != p.den)
        for (double &v : values) v = v * (double)p.num / (double)p.den;
    return values;
}

}  // namespace

int main(int argc, char **argv) {
    const char *file = nullptr, *runner = nullptr, *kernel = "f64", *examples = nullptr, *write = nullptr;
    size_t tile = 0;
    bool host_reduce = false, dump = false, keyed = false;
    for (int k = 1; k < argc; k++) {
        std::string a = argv[k];
        if (a == "--runner" && k + 1 < argc) runner = argv[++k];
        else if (a == "--kernel" && k + 1 < argc) kernel = argv[++k];
        else if (a == "--tile" && k + 1 < argc) tile = strtoul(argv[++k], nullptr, 10);
        else if (a == "--host-reduce") host_reduce = true;
        else if (a == "--key") keyed = true;
        else if (a == "--dump") dump = true;
        else if (a == "--examples" && k + 1 < argc) examples = argv[++k];
        else if (a == "--write-examples" && k + 1 < argc) write = argv[++k];
        else if (!file) file = argv[k];
        else {
            fprintf(stderr, "pacuda-check: unknown argument %s\n", argv[k]);
            return 2;
        }
    }
    if (!file || (!runner && !dump && !write)) {
        fprintf(stderr, "usage: pacuda-check FILE.pacuda --runner LIBRARY [--kernel f64|df64|mixed|mixed64] [--tile N] "
                        "[--host-reduce] [--key] [--dump] [--examples EXAMPLES.json]\n"
                        "       pacuda-check FILE.pacuda --write-examples EXAMPLES.json\n");
        return 2;
    }
    try {
        FILE *fp = fopen(file, "rb");
        if (!fp) throw std::runtime_error(std::string(file) + ": cannot open");
        std::vector<uint8_t> bytes;
        uint8_t buf[1 << 16];
        size_t got;
        while ((got = fread(buf, 1, sizeof buf, fp)) > 0) bytes.insert(bytes.end(), buf, buf + got);
        fclose(fp);
        File f = read_file(bytes);
        if (write) {
            write_examples(f, write);
            if (!runner && !dump) return 0;
        }
        if (examples) {
            const std::string text = read_text(examples);
            take_examples(f, JsonParser(text).document());
        }
        if (dump) {
            printf("kernel %u batch %u symjit %zu processes %zu\n", f.kernel, f.batch, f.sha256.size(), f.processes.size());
            for (const Process &p : f.processes)
                for (const Colour &c : p.colours)
                    printf("%s %s legs %u helicities %u words %zu planes %u/%u/%u points %u\n", p.key.c_str(), c.id.c_str(),
                           p.legs, p.helicities, c.words.size(), c.current_planes, c.amplitude_planes, c.momentum_planes,
                           c.points);
            if (!runner) return 0;
        }
        void *lib = dlopen(runner, RTLD_NOW);
        if (!lib) throw std::runtime_error(dlerror());
        Runner rn;
        rn.run = (int (*)(const pacman_recurrence_call *))dlsym(lib, "pacman_recurrence_v1_run");
        rn.reduce_resolved = (int (*)(const pacman_recurrence_reduce_resolved *))dlsym(lib, "pacman_recurrence_v1_reduce_resolved");
        rn.configure = (int (*)(const pacman_recurrence_options *))dlsym(lib, "pacman_recurrence_v1_configure");
        rn.symjit_kernel = (int (*)(const char *))dlsym(lib, "pacman_recurrence_v1_symjit_kernel");
        if (!rn.run || !rn.reduce_resolved || !rn.configure) throw std::runtime_error("the library lacks the v1 entry points");
        pacman_recurrence_options o;
        memset(&o, 0, sizeof o);
        o.abi_version = PACMAN_RECURRENCE_ABI;
        o.struct_size = sizeof o;
        std::string k = kernel;
        o.kernel = k == "f64" ? PACMAN_RECURRENCE_KERNEL_F64 : k == "mixed" ? PACMAN_RECURRENCE_KERNEL_MIXED
                 : k == "mixed64" ? PACMAN_RECURRENCE_KERNEL_MIXED64 : PACMAN_RECURRENCE_KERNEL_DF64;
        if (rn.configure(&o) != 0) throw std::runtime_error("configure failed");
        std::vector<uint32_t> map;
        for (const std::string &s : f.sha256) {
            int idx = rn.symjit_kernel ? rn.symjit_kernel(s.c_str()) : -1;
            if (idx < 0) throw std::runtime_error("the runner has no symjit kernel " + s);
            map.push_back((uint32_t)idx);
        }
        size_t total_differ = 0, total_values = 0;
        for (const Process &p : f.processes)
            for (const Colour &c : p.colours) {
                size_t t = tile ? tile : (f.batch ? f.batch : c.tile);
                std::vector<double> got = evaluate(rn, map, p, c, t, host_reduce, keyed);
                size_t differ = 0;
                double worst = 0, worst_total = 0;
                const size_t h = std::max<uint32_t>(1, p.helicities);
                for (size_t q = 0; q < c.points; q++) {
                    double scale = 0, ta = 0, tb = 0;
                    for (size_t o = 0; o < h; o++) {
                        double a = got[q * h + o], b = c.values[q * h + o];
                        if (memcmp(&a, &b, 8)) differ++;
                        scale += fabs(b);
                        ta += a;
                        tb += b;
                    }
                    if (scale == 0) continue;
                    for (size_t o = 0; o < h; o++) worst = fmax(worst, fabs(got[q * h + o] - c.values[q * h + o]) / scale);
                    worst_total = fmax(worst_total, fabs(ta / tb - 1.0));
                }
                printf("%s %s: %zu points, %zu of %zu values differ in bits, worst %.3e of the point's total, worst "
                       "total %.3e\n", p.key.c_str(), c.id.c_str(), (size_t)c.points, differ, got.size(), worst, worst_total);
                total_differ += differ;
                total_values += got.size();
            }
        // (a library that runs some calls on other kernels says how many: this library's fused ones)
        if (auto fused = (uint64_t (*)(void))dlsym(lib, "pacuda_v1_fused_runs"))
            printf("runs on the fused kernels: %llu\n", (unsigned long long)fused());
        if (k == "f64" && total_differ) {
            fprintf(stderr, "pacuda-check: %zu of %zu f64 values differ from the recording\n", total_differ, total_values);
            return 1;
        }
        return 0;
    } catch (const std::exception &e) {
        fprintf(stderr, "pacuda-check: %s\n", e.what());
        return 1;
    }
}
