// Fuzz driver for the sanitizer build of the engine's host code.  Built with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/sanitize/Makefile); any report aborts the process.
//
//   fuzz_host <seed files...>
//
// A seed file holds one protobuf message, its kind given by its name's prefix: "ckt" (spice21.Circuit),
// "op" (spice21.Op), "tran" (spice21.Tran) or "ac" (spice21.Ac).  For every seed:
//   * the seed itself: decodes, elaborates, and runs the host symbolic phase (element creation, the
//     first-iteration assembly at x = 0 and at random voltages, the pivot search, schedule and records);
//   * every truncation, and a few thousand mutations (byte overwrites, insertions, deletions, splices
//     of two seeds): each must decode, or return S21_DECODE (bytes that are no Circuit) or S21_ERR with
//     the reference's conversion error (a Circuit naming no component variant, circuit.rs); what decodes
//     goes through elaboration and, when that succeeds, the symbolic phase, whose errors must be S21_*
//     codes;
//   * the byte ABI entry (s21_op_bytes / s21_tran_bytes / s21_ac_bytes) on the seed and its mutations:
//     S21_DECODE for bytes that do not decode, else an elaboration error or the stub engine's S21_DEVICE.
// Then random buffers (the same outcomes), and random sparse systems through the
// single-large-instance symbolic phase and its CPU numeric reference.  Prints a summary; exit 0 on success.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"
#include "large.h"

namespace {

uint64_t g_rng = 0x9e3779b97f4a7c15ull;
uint64_t next_u64() {  // splitmix64
    uint64_t z = (g_rng += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
size_t below(size_t n) { return n ? (size_t)(next_u64() % n) : 0; }
double uniform() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

struct Counts {
    long decoded = 0, rejected = 0, elaborated = 0, elab_errors = 0, symbolic = 0, sym_errors = 0;
    long abi_calls = 0, large_systems = 0, large_singular = 0;
} C;

int fail_count = 0;
void check(bool ok, const char* what, int code) {
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s (code %d)\n", what, code);
        fail_count++;
    }
}

// The host symbolic phase of one flattened netlist, optionally for a transient with an IC on var 0.
void symbolic(const s21_netlist& nl, bool tran) {
    using namespace s21;
    try {
        std::vector<std::string> names(nl.n_vars);
        for (int k = 0; k < nl.n_vars; k++) names[k] = nl.var_name && nl.var_name[k] ? nl.var_name[k] : "";
        std::vector<int> kinds(nl.var_kind, nl.var_kind + nl.n_vars);
        std::vector<s21_comp> comps(nl.comps, nl.comps + nl.n_comps);
        int32_t ic_var = 0;
        double ic_val = 0.25;
        const char* ic_name = nl.n_vars > 0 && nl.var_name ? nl.var_name[0] : "";
        s21_tran_options tr{1e-12, 1e-11, 1, &ic_var, &ic_val, &ic_name};
        Topology t;
        build_comps(nl.n_vars, names, kinds, comps, nl.n_mos1_models, nl.n_mos1_insts, nl.n_diode_models,
                    nl.n_diode_insts, tran && nl.n_vars > 0 ? &tr : nullptr, t);
        topo_elements(t);
        s21_options o;
        s21_default_options(&o);
        HostInstance h;
        std::vector<uint8_t> caps;
        bool bsim4 = false;
        for (const Comp& c : t.comps) {
            if (c.type == S21_COMP_BSIM4) {  // single-instance path only: derivation and assembly, no schedule
                try {
                    h.b4.push_back(b4::derive_params(nl.bsim4_models[c.model], nl.bsim4_insts[c.inst]));
                } catch (const std::domain_error&) {
                    fail(S21_ERR, "BSIM4 derivation");
                }
                bsim4 = true;
            } else if (c.type == S21_COMP_MOS1) {
                h.m1.push_back(derive_mos1(nl.mos1_models[c.model], nl.mos1_insts[c.inst], o));
                caps.push_back(h.m1.back().has_caps ? 1 : 0);
            } else if (c.type == S21_COMP_DIODE) {
                h.dio.push_back(derive_diode(nl.diode_models[c.model], nl.diode_insts[c.inst], o));
            } else if (c.type != S21_COMP_MOS0) {
                h.lin_value.push_back(c.value);
            }
        }
        std::vector<double> x(t.n, 0.0), vals, rhs;
        host_assemble_op(t, h, o, x, vals, rhs);
        topo_pivot(t, vals);
        for (double& v : x) v = 2.0 * uniform() - 0.5;
        host_assemble_op(t, h, o, x, vals, rhs);
        if (t.sym_status == S21_OK && !bsim4) {
            topo_schedule(t, caps);
            topo_records(t);
        }
        C.symbolic++;
    } catch (const s21::Error& e) {
        check(e.code > 0 && e.code < 16, "symbolic phase error code", e.code);
        C.sym_errors++;
    }
}

// Decode one candidate Circuit; on success elaborate and run the symbolic phase.
void try_circuit(const std::vector<uint8_t>& b, bool must_decode) {
    s21_ckt* ckt = nullptr;
    const int rc = s21_ckt_decode(b.data(), b.size(), &ckt);
    if (rc != S21_OK) {
        // S21_DECODE for bytes that are no Circuit; S21_ERR with the reference's text for a Circuit whose
        // conversion fails (an Instance or Def with no variant set: "Invalid Component", circuit.rs)
        char msg[256] = {0};
        s21_last_error(msg, sizeof msg);
        if (rc != S21_DECODE && std::getenv("FUZZ_VERBOSE")) std::fprintf(stderr, "decode code %d: %s\n", rc, msg);
        check(rc == S21_DECODE || (rc == S21_ERR && msg[0]), "decode returns S21_DECODE or a conversion error", rc);
        check(!must_decode, "seed decodes", rc);
        check(ckt == nullptr, "no circuit on a decode error", rc);
        C.rejected++;
        return;
    }
    C.decoded++;
    s21_elab* e = nullptr;
    const int re = s21_elaborate(ckt, &e);
    if (re == S21_OK) {
        C.elaborated++;
        s21_netlist nl;
        const int rn = s21_elab_netlist(e, &nl);
        check(rn == S21_OK, "netlist view", rn);
        if (rn == S21_OK && nl.n_vars < 4000) {
            symbolic(nl, false);
            symbolic(nl, true);
        }
        s21_elab_free(e);
    } else {
        check(re > 0 && re < 16 && re != S21_DECODE, "elaboration error code", re);
        C.elab_errors++;
    }
    s21_ckt_free(ckt);
}

using AbiFn = int (*)(const uint8_t*, size_t, uint8_t**, size_t*);

void try_abi(AbiFn fn, const std::vector<uint8_t>& b) {
    uint8_t* out = nullptr;
    size_t n = 0;
    const int rc = fn(b.data(), b.size(), &out, &n);
    check(rc != S21_OK, "the stub engine cannot succeed", rc);
    check(rc > 0 && rc < 16, "byte ABI returns an S21_* code", rc);
    check(out == nullptr, "no output on error", rc);
    C.abi_calls++;
}

std::vector<uint8_t> mutate(const std::vector<uint8_t>& s, const std::vector<std::vector<uint8_t>>& all) {
    std::vector<uint8_t> b = s;
    const int edits = 1 + (int)below(4);
    for (int k = 0; k < edits; k++) {
        switch (below(5)) {
            case 0:  // overwrite a byte
                if (!b.empty()) b[below(b.size())] = (uint8_t)next_u64();
                break;
            case 1:  // flip a bit
                if (!b.empty()) b[below(b.size())] ^= (uint8_t)(1u << below(8));
                break;
            case 2: {  // insert bytes (often varint continuation bytes)
                const size_t at = below(b.size() + 1), n = 1 + below(6);
                for (size_t i = 0; i < n; i++) b.insert(b.begin() + at, below(2) ? 0xff : (uint8_t)next_u64());
                break;
            }
            case 3: {  // delete a range
                if (b.empty()) break;
                const size_t at = below(b.size()), n = 1 + below(std::min<size_t>(8, b.size() - at));
                b.erase(b.begin() + at, b.begin() + at + n);
                break;
            }
            default: {  // splice a range of another seed
                const std::vector<uint8_t>& o = all[below(all.size())];
                if (o.empty()) break;
                const size_t from = below(o.size()), n = 1 + below(std::min<size_t>(32, o.size() - from));
                const size_t at = below(b.size() + 1);
                b.insert(b.begin() + at, o.begin() + from, o.begin() + from + n);
                break;
            }
        }
    }
    return b;
}

void large_systems(int count) {
    for (int s = 0; s < count; s++) {
        const int n = 1 + (int)below(150);
        std::vector<int> r, c;
        std::vector<double> v;
        const bool singular = below(8) == 0;
        for (int i = 0; i < n; i++) {
            if (singular && i == n / 2) continue;  // an empty row and column
            r.push_back(i);
            c.push_back(i);
            v.push_back(4.0 + uniform());
        }
        const int extra = (int)below(4 * n + 1);
        for (int k = 0; k < extra; k++) {
            const int i = (int)below(n), j = (int)below(n);
            if (singular && (i == n / 2 || j == n / 2)) continue;
            r.push_back(i);
            c.push_back(j);
            v.push_back(uniform() - 0.5);
        }
        try {
            s21::LargeSym S;
            s21::large_symbolic(n, r, c, &v, S, below(2) ? 0.0 : 0.3);
            std::vector<double> pool, x, b(n);
            for (double& y : b) y = uniform();
            s21::large_factor_host(S, r, c, v, pool);
            s21::large_solve_host(S, pool, b, x);
            check((int)x.size() == n, "solution length", 0);
            check(!singular, "a system with an empty row and column is reported singular", 0);
            C.large_systems++;
        } catch (const s21::Error& e) {
            check(singular && e.code > 0 && e.code < 16, "only the singular systems fail", e.code);
            C.large_singular++;
        }
    }
}

std::vector<uint8_t> read_file(const char* path) {
    std::vector<uint8_t> b;
    FILE* f = std::fopen(path, "rb");
    if (!f) return b;
    uint8_t buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + n);
    std::fclose(f);
    return b;
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<std::vector<uint8_t>> seeds;
    std::vector<std::string> kinds;
    for (int a = 1; a < argc; a++) {
        const std::string path = argv[a];
        const std::string base = path.substr(path.find_last_of('/') + 1);
        seeds.push_back(read_file(argv[a]));
        kinds.push_back(base.substr(0, base.find_first_of("_.")));
    }
    const char* env = std::getenv("FUZZ_MUTATIONS");
    const int mutations = env ? std::atoi(env) : 20000;
    for (size_t s = 0; s < seeds.size(); s++) {
        const std::vector<uint8_t>& b = seeds[s];
        const std::string& k = kinds[s];
        if (k == "ckt") {
            try_circuit(b, true);
            for (size_t len = 0; len < b.size(); len++) try_circuit(std::vector<uint8_t>(b.begin(), b.begin() + len), false);
            for (int m = 0; m < mutations; m++) try_circuit(mutate(b, seeds), false);
        } else {
            AbiFn fn = k == "op" ? s21_op_bytes : k == "tran" ? s21_tran_bytes : s21_ac_bytes;
            try_abi(fn, b);
            for (size_t len = 0; len < b.size(); len++) try_abi(fn, std::vector<uint8_t>(b.begin(), b.begin() + len));
            for (int m = 0; m < mutations / 3; m++) try_abi(fn, mutate(b, seeds));
        }
    }
    for (int r = 0; r < mutations; r++) {
        std::vector<uint8_t> b(below(300));
        for (uint8_t& x : b) x = (uint8_t)next_u64();
        try_circuit(b, false);
    }
    large_systems(400);
    std::printf("decoded %ld rejected %ld elaborated %ld elab_errors %ld symbolic %ld sym_errors %ld abi %ld large %ld "
                "large_singular %ld failures %d\n",
                C.decoded, C.rejected, C.elaborated, C.elab_errors, C.symbolic, C.sym_errors, C.abi_calls,
                C.large_systems, C.large_singular, fail_count);
    return fail_count ? 1 : 0;
}
