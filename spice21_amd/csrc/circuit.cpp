// Circuits: protobuf decoding, hierarchy elaboration and the byte ABI.
//
//  * a proto3 wire-format reader/writer for the subset of protos/spice21.proto and protos/mos.proto
//    the analyses take and return (Circuit, Op/Tran/Ac, SimOptions, TranOptions, AcOptions, OpResult,
//    TranResult, AcResult);
//  * Ckt::from_proto (circuit.rs:281-330) and the Elaborator (elab.rs:22-265), producing the flattened
//    netlist the engine runs;
//  * CallableProto::call_bytes for Op, Tran and Ac (proto.rs:32-145).
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "internal.h"

namespace s21 {

// ============================================================================================
// proto3 wire format
// ============================================================================================
struct Reader {
    const uint8_t* p;
    const uint8_t* end;
    Reader(const uint8_t* b, size_t n) : p(b), end(b + n) {}
    bool done() const { return p >= end; }
    [[noreturn]] void bad(const char* what) { fail(S21_DECODE, std::string("Decode Error: ") + what); }
    uint64_t varint() {
        uint64_t v = 0;
        for (int shift = 0; shift < 64; shift += 7) {
            if (p >= end) bad("buffer underflow");
            uint8_t b = *p++;
            v |= (uint64_t)(b & 0x7f) << shift;
            if (!(b & 0x80)) return v;
        }
        bad("invalid varint");
    }
    double f64() {
        if (end - p < 8) bad("buffer underflow");
        double d;
        memcpy(&d, p, 8);
        p += 8;
        return d;
    }
    Reader sub() {
        uint64_t n = varint();
        if ((uint64_t)(end - p) < n) bad("buffer underflow");
        Reader r(p, (size_t)n);
        p += n;
        return r;
    }
    std::string str() {
        Reader r = sub();
        return std::string((const char*)r.p, (const char*)r.end);
    }
    void skip(int wt) {
        switch (wt) {
        case 0: varint(); break;
        case 1:
            if (end - p < 8) bad("buffer underflow");
            p += 8;
            break;
        case 2: sub(); break;
        case 5:
            if (end - p < 4) bad("buffer underflow");
            p += 4;
            break;
        default: bad("invalid wire type");
        }
    }
    // Next (field, wire type); false at end.
    bool next(int& field, int& wt) {
        if (done()) return false;
        uint64_t k = varint();
        field = (int)(k >> 3);
        wt = (int)(k & 7);
        if (field == 0) bad("invalid tag");
        return true;
    }
};

// google.protobuf.DoubleValue / Int64Value
static double wrapped_double(Reader r) {
    double v = 0.0;
    int f, wt;
    while (r.next(f, wt)) {
        if (f == 1 && wt == 1)
            v = r.f64();
        else
            r.skip(wt);
    }
    return v;
}
static double wrapped_int64(Reader r) {
    int64_t v = 0;
    int f, wt;
    while (r.next(f, wt)) {
        if (f == 1 && wt == 0)
            v = (int64_t)r.varint();
        else
            r.skip(wt);
    }
    return (double)v;
}

struct Writer {
    std::vector<uint8_t> b;
    void varint(uint64_t v) {
        while (v >= 0x80) {
            b.push_back((uint8_t)(v | 0x80));
            v >>= 7;
        }
        b.push_back((uint8_t)v);
    }
    void tag(int f, int wt) { varint(((uint64_t)f << 3) | (uint64_t)wt); }
    void f64(int f, double d) {
        tag(f, 1);
        uint8_t t[8];
        memcpy(t, &d, 8);
        b.insert(b.end(), t, t + 8);
    }
    void bytes(int f, const std::vector<uint8_t>& m) {
        tag(f, 2);
        varint(m.size());
        b.insert(b.end(), m.begin(), m.end());
    }
    void str(int f, const std::string& s) {
        tag(f, 2);
        varint(s.size());
        b.insert(b.end(), s.begin(), s.end());
    }
    void packed_f64(int f, const double* v, size_t n) {
        tag(f, 2);
        varint(n * 8);
        const uint8_t* q = (const uint8_t*)v;
        b.insert(b.end(), q, q + n * 8);
    }
};

// ============================================================================================
// Circuit model (circuit.rs)
// ============================================================================================
struct CompSpec {
    enum Kind { R, C, I, V, D, M, X } kind;
    std::string name, p, n;
    double value = 0.0, acm = 0.0;
    std::string model, params;          // Mos, Diode
    std::string d, g, s, b;             // Mos ports
    std::string module;                 // module instance
    std::vector<std::pair<std::string, std::string>> ports;
};
struct ModuleSpec {
    std::string name;
    std::vector<std::string> ports, signals;
    std::vector<CompSpec> comps;
};

}  // namespace s21

struct s21_ckt {
    std::string name;
    std::vector<std::string> signals;
    std::vector<s21::CompSpec> comps;
    std::map<std::string, s21_params> mos1_models, mos1_insts, diode_models, diode_insts;
    std::map<std::string, int> mos0;
    std::map<std::string, s21_params> bsim4_models, bsim4_insts;
    std::map<std::string, s21::ModuleSpec> modules;
};

namespace s21 {

static s21_params params_msg(Reader r, int int64_field) {
    s21_params p;
    memset(&p, 0, sizeof p);
    int f, wt;
    while (r.next(f, wt)) {
        if (f >= 2 && f < S21_PARAM_SLOTS && wt == 2) {
            Reader sub = r.sub();
            p.v[f] = (f == int64_field) ? wrapped_int64(sub) : wrapped_double(sub);
            p.set |= 1ull << f;
        } else if (f == 2 && wt == 0) {  // Mos1Model.mos_type (enum)
            p.v[2] = (double)r.varint();
            p.set |= 1ull << 2;
        } else {
            r.skip(wt);
        }
    }
    return p;
}
// Bsim4Model (protos/bsim4.proto:44-46, name 900): only the polarity reaches the model
// (Bsim4ModelSpecs::from, comps/bsim4/model/mod.rs:21-30; mos_type 1 is PMOS, anything else NMOS).
static s21_params bsim4_model_msg(Reader r, std::string& name) {
    s21_params p;
    memset(&p, 0, sizeof p);
    int f, wt;
    while (r.next(f, wt)) {
        if (f == 1 && wt == 0) {
            p.v[1] = r.varint() == 1 ? 1.0 : 0.0;
            p.set |= 1ull << 1;
        } else if (f == 900 && wt == 2) {
            name = r.str();
        } else {
            r.skip(wt);
        }
    }
    return p;
}
// Bsim4InstParams (protos/bsim4.proto:6-42): fields 1-31 are DoubleValue / UInt64Value wrappers, name 40.
static s21_params bsim4_inst_msg(Reader r, std::string& name) {
    s21_params p;
    memset(&p, 0, sizeof p);
    int f, wt;
    while (r.next(f, wt)) {
        if (f >= 1 && f <= 31 && wt == 2) {
            Reader sub = r.sub();
            const bool integer = f == 17 || f == 18 || f >= 27;
            p.v[f] = integer ? (double)wrapped_int64(sub) : wrapped_double(sub);
            p.set |= 1ull << f;
        } else if (f == 40 && wt == 2) {
            name = r.str();
        } else {
            r.skip(wt);
        }
    }
    return p;
}
static std::string name_of(Reader r) {
    int f, wt;
    std::string nm;
    while (r.next(f, wt)) {
        if (f == 1 && wt == 2)
            nm = r.str();
        else
            r.skip(wt);
    }
    return nm;
}
static void string_map_entry(Reader r, std::string& k, std::string& v) {
    int f, wt;
    while (r.next(f, wt)) {
        if (f == 1 && wt == 2)
            k = r.str();
        else if (f == 2 && wt == 2)
            v = r.str();
        else
            r.skip(wt);
    }
}
static void double_map_entry(Reader r, std::string& k, double& v) {
    int f, wt;
    while (r.next(f, wt)) {
        if (f == 1 && wt == 2)
            k = r.str();
        else if (f == 2 && wt == 1)
            v = r.f64();
        else
            r.skip(wt);
    }
}

// Two-terminal element: name 1, p 2, n 3, value 4 (, acm 5).
static void two_term(Reader r, CompSpec& c) {
    int f, wt;
    while (r.next(f, wt)) {
        if (f == 1 && wt == 2)
            c.name = r.str();
        else if (f == 2 && wt == 2)
            c.p = r.str();
        else if (f == 3 && wt == 2)
            c.n = r.str();
        else if (f == 4 && wt == 1)
            c.value = r.f64();
        else if (f == 5 && wt == 1)
            c.acm = r.f64();
        else
            r.skip(wt);
    }
}

// Instance (protos/spice21.proto:80-90) -> circuit::Comp (circuit.rs:150-215).  Returns false for an
// Instance with no component set.
static bool instance(Reader r, CompSpec& c) {
    int f, wt;
    bool any = false;
    while (r.next(f, wt)) {
        if (wt != 2 || f < 1 || f > 7) {
            r.skip(wt);
            continue;
        }
        Reader m = r.sub();
        any = true;
        c = CompSpec{};
        switch (f) {
        case 1: c.kind = CompSpec::R; two_term(m, c); break;
        case 2: c.kind = CompSpec::C; two_term(m, c); break;
        case 4: c.kind = CompSpec::I; two_term(m, c); break;
        case 5: c.kind = CompSpec::V; two_term(m, c); break;
        case 3: {  // Mos: name 1, model 2, params 3, ports 4 (MosPorts d 1, g 2, s 3, b 4)
            c.kind = CompSpec::M;
            int g2, w2;
            while (m.next(g2, w2)) {
                if (g2 == 1 && w2 == 2)
                    c.name = m.str();
                else if (g2 == 2 && w2 == 2)
                    c.model = m.str();
                else if (g2 == 3 && w2 == 2)
                    c.params = m.str();
                else if (g2 == 4 && w2 == 2) {
                    Reader pr = m.sub();
                    int g3, w3;
                    while (pr.next(g3, w3)) {
                        if (w3 != 2) {
                            pr.skip(w3);
                            continue;
                        }
                        std::string v = pr.str();
                        if (g3 == 1) c.d = v;
                        else if (g3 == 2) c.g = v;
                        else if (g3 == 3) c.s = v;
                        else if (g3 == 4) c.b = v;
                    }
                } else
                    m.skip(w2);
            }
            break;
        }
        case 6: {  // Diode: name 1, p 2, n 3, model 4, params 5
            c.kind = CompSpec::D;
            int g2, w2;
            while (m.next(g2, w2)) {
                if (w2 != 2) {
                    m.skip(w2);
                    continue;
                }
                std::string v = m.str();
                if (g2 == 1) c.name = v;
                else if (g2 == 2) c.p = v;
                else if (g2 == 3) c.n = v;
                else if (g2 == 4) c.model = v;
                else if (g2 == 5) c.params = v;
            }
            break;
        }
        case 7: {  // ModuleInstance: name 1, module 2, ports 3 (map<string,string>), params 4
            c.kind = CompSpec::X;
            int g2, w2;
            while (m.next(g2, w2)) {
                if (g2 == 1 && w2 == 2)
                    c.name = m.str();
                else if (g2 == 2 && w2 == 2)
                    c.module = m.str();
                else if (g2 == 3 && w2 == 2) {
                    std::string k, v;
                    string_map_entry(m.sub(), k, v);
                    c.ports.push_back({k, v});
                } else
                    m.skip(w2);
            }
            break;
        }
        }
    }
    return any;
}

static void module_def(Reader r, ModuleSpec& md) {
    int f, wt;
    while (r.next(f, wt)) {
        if (f == 1 && wt == 2)
            md.name = r.str();
        else if (f == 2 && wt == 2)
            md.ports.push_back(r.str());
        else if (f == 4 && wt == 2)
            md.signals.push_back(r.str());
        else if (f == 5 && wt == 2) {
            CompSpec c;
            if (!instance(r.sub(), c)) fail(S21_ERR, "Invalid Comp!!!");  // elab.rs:232-236
            md.comps.push_back(c);
        } else
            r.skip(wt);
    }
}

// Circuit (protos/spice21.proto:128-133) -> Ckt (circuit.rs:281-330).
static void decode_circuit(Reader r, s21_ckt& ck) {
    int f, wt;
    while (r.next(f, wt)) {
        if (f == 1 && wt == 2)
            ck.name = r.str();
        else if (f == 2 && wt == 2)
            ck.signals.push_back(r.str());
        else if (f == 3 && wt == 2) {  // Def
            Reader d = r.sub();
            int g, w;
            bool any = false;
            while (d.next(g, w)) {
                if (w != 2) {
                    d.skip(w);
                    continue;
                }
                Reader m = d.sub();
                any = true;
                switch (g) {
                case 1: {
                    ModuleSpec md;
                    module_def(m, md);
                    ck.modules[md.name] = md;
                    break;
                }
                case 2: ck.diode_models[name_of(m)] = params_msg(m, -1); break;
                case 3: ck.diode_insts[name_of(m)] = params_msg(m, -1); break;
                case 4: {  // Defs::add_model keyed by name (circuit.rs:297-301)
                    std::string nm;
                    s21_params p = bsim4_model_msg(m, nm);
                    ck.bsim4_models[nm] = p;
                    break;
                }
                case 5: {  // Bsim4Cache::add_inst keyed by name (cache.rs:55-57)
                    std::string nm;
                    s21_params p = bsim4_inst_msg(m, nm);
                    ck.bsim4_insts[nm] = p;
                    break;
                }
                case 6: ck.mos1_models[name_of(m)] = params_msg(m, 28); break;
                case 7: ck.mos1_insts[name_of(m)] = params_msg(m, -1); break;
                default: any = false; break;
                }
            }
            if (!any) fail(S21_ERR, "Invalid Definition");  // `def.defines.unwrap()`, circuit.rs:293
        } else if (f == 4 && wt == 2) {
            CompSpec c;
            if (!instance(r.sub(), c)) fail(S21_ERR, "Invalid Component");  // circuit.rs:322-327
            ck.comps.push_back(c);
        } else
            r.skip(wt);
    }
}

// ============================================================================================
// Elaboration (elab.rs)
// ============================================================================================
}  // namespace s21

struct s21_elab {
    std::vector<std::string> names;
    std::vector<int32_t> kinds;
    std::vector<const char*> name_ptrs;
    std::vector<s21_comp> comps;
    std::vector<s21_params> mos1_models, mos1_insts, diode_models, diode_insts, bsim4_models, bsim4_insts;
    std::vector<std::string> comp_names;
};

namespace s21 {

struct Elaborator {
    const s21_ckt& ck;
    s21_elab& e;
    std::vector<std::string> path;
    std::unordered_map<std::string, int> first;  // name -> first variable with that name
    std::map<std::string, int> m1_model_idx, m1_inst_idx, d_model_idx, d_inst_idx, b4_model_idx, b4_inst_idx;
    using Ns = std::map<std::string, int>;  // namespace: node name -> variable (S21_GND for ground)

    Elaborator(const s21_ckt& c, s21_elab& out) : ck(c), e(out) {}

    int add(const std::string& name, int kind) {  // Variables::add, analysis.rs:67-73
        e.names.push_back(name);
        e.kinds.push_back(kind);
        int k = (int)e.names.size() - 1;
        first.emplace(name, k);
        return k;
    }
    int find_or_create(const std::string& name) {  // analysis.rs:92-111
        auto it = first.find(name);
        if (it != first.end()) return it->second;
        return add(name, S21_VAR_V);
    }
    std::string pathstr() const {
        std::string s;
        for (size_t k = 0; k < path.size(); k++) {
            if (k) s += ".";
            s += path[k];
        }
        return s;
    }
    bool on_top() const { return path.empty(); }
    int node_var(const std::string& node, bool autonode, Ns& ns) {  // elab.rs:35-52
        if (autonode) {
            if (node.empty()) return S21_GND;
            path.push_back(node);
            int v = find_or_create(pathstr());
            ns[node] = v;
            path.pop_back();
            return v;
        }
        auto it = ns.find(node);
        if (it == ns.end()) fail(S21_ERR, "!!!");
        return it->second;
    }
    void signal(const std::string& sig, Ns& ns) {  // elab.rs:209-216
        path.push_back(sig);
        int v = add(pathstr(), S21_VAR_V);
        ns[sig] = v;
        path.pop_back();
    }
    template <typename M>
    int def_index(const std::string& key, const M& table, std::map<std::string, int>& idx,
                  std::vector<s21_params>& out) {
        auto it = idx.find(key);
        if (it != idx.end()) return it->second;
        out.push_back(table.at(key));
        idx[key] = (int)out.size() - 1;
        return (int)out.size() - 1;
    }
    void push(int type, const int* nd, int nn, int model, int inst, double value, double acm, const std::string& nm) {
        s21_comp c;
        memset(&c, 0, sizeof c);
        c.type = type;
        for (int k = 0; k < 6; k++) c.node[k] = k < nn ? nd[k] : S21_GND;
        c.model = model;
        c.inst = inst;
        c.value = value;
        c.acm = acm;
        e.comps.push_back(c);
        e.comp_names.push_back(nm);
    }
    void instance(const CompSpec& c, Ns& ns, bool autonode) {  // elab.rs:55-84
        switch (c.kind) {
        case CompSpec::R:
        case CompSpec::C:
        case CompSpec::I: {
            int nd[2] = {node_var(c.p, autonode, ns), node_var(c.n, autonode, ns)};
            const int type = c.kind == CompSpec::R ? S21_COMP_R : c.kind == CompSpec::C ? S21_COMP_C : S21_COMP_I;
            push(type, nd, 2, 0, 0, c.value, 0.0, c.name);
            break;
        }
        case CompSpec::V: {  // elab.rs:112-127
            int p = node_var(c.p, on_top(), ns), n = node_var(c.n, on_top(), ns);
            path.push_back(c.name);
            int iv = add(pathstr(), S21_VAR_I);
            path.pop_back();
            int nd[3] = {p, n, iv};
            push(S21_COMP_V, nd, 3, 0, 0, c.value, c.acm, c.name);
            break;
        }
        case CompSpec::D: {  // elab.rs:85-111
            auto pi = ns.find(c.p), ni = ns.find(c.n);
            if (pi == ns.end() || ni == ns.end()) fail(S21_ERR, "Diode node not defined");
            path.push_back(c.name);
            if (!ck.diode_insts.count(c.params) || !ck.diode_models.count(c.model))
                fail(S21_ERR, "Parameters not defined: " + c.params);
            const int mi = def_index(c.model, ck.diode_models, d_model_idx, e.diode_models);
            const int ii = def_index(c.params, ck.diode_insts, d_inst_idx, e.diode_insts);
            const int r = diode_model_has_rs(ck.diode_models.at(c.model)) ? add(pathstr() + ".r", S21_VAR_V) : pi->second;
            path.pop_back();
            int nd[3] = {pi->second, ni->second, r};
            push(S21_COMP_DIODE, nd, 3, mi, ii, 0.0, 0.0, c.name);
            break;
        }
        case CompSpec::M: {  // elab.rs:128-173
            int nd[6] = {node_var(c.d, on_top(), ns), node_var(c.g, on_top(), ns), node_var(c.s, on_top(), ns),
                         node_var(c.b, on_top(), ns), 0, 0};
            path.push_back(c.name);
            if (ck.bsim4_models.count(c.model)) {  // BSIM4 definitions take precedence (elab.rs:143-147)
                if (!ck.bsim4_insts.count(c.params)) fail(S21_ERR, "Parameters not defined: " + c.params);
                const int mi = def_index(c.model, ck.bsim4_models, b4_model_idx, e.bsim4_models);
                const int ii = def_index(c.params, ck.bsim4_insts, b4_inst_idx, e.bsim4_insts);
                path.pop_back();  // no internal nodes at the default mode selectors (bsim4ports.rs:31-112)
                push(S21_COMP_BSIM4, nd, 4, mi, ii, 0.0, 0.0, c.name);
            } else if (ck.mos1_models.count(c.model)) {
                if (!ck.mos1_insts.count(c.params)) fail(S21_ERR, "Parameters not defined: " + c.params);
                const s21_params& mp = ck.mos1_models.at(c.model);
                const int mi = def_index(c.model, ck.mos1_models, m1_model_idx, e.mos1_models);
                const int ii = def_index(c.params, ck.mos1_insts, m1_inst_idx, e.mos1_insts);
                const bool rd = ((mp.set >> 8) & 1) || ((mp.set >> 17) & 1);
                const bool rs = ((mp.set >> 9) & 1) || ((mp.set >> 17) & 1);
                nd[4] = rd ? add(pathstr() + ".dp", S21_VAR_V) : nd[0];  // comps/mos.rs:599-610
                nd[5] = rs ? add(pathstr() + ".sp", S21_VAR_V) : nd[2];
                path.pop_back();
                push(S21_COMP_MOS1, nd, 6, mi, ii, 0.0, 0.0, c.name);
            } else if (ck.mos0.count(c.model)) {
                path.pop_back();
                push(S21_COMP_MOS0, nd, 4, ck.mos0.at(c.model), 0, 0.0, 0.0, c.name);
            } else {
                fail(S21_ERR, "Model not defined: " + c.model);
            }
            break;
        }
        case CompSpec::X: {  // elab.rs:182-207
            auto it = ck.modules.find(c.module);
            if (it == ck.modules.end()) fail(S21_ERR, "ModuleDef not found: " + c.module);
            Ns inst_ns;
            for (auto& kv : c.ports) {
                auto v = ns.find(kv.second);
                if (v == ns.end()) fail(S21_ERR, "Module port connected to an undefined signal: " + kv.second);
                inst_ns[kv.first] = v->second;
            }
            path.push_back(c.name);
            if (path.size() > 1024) fail(S21_ERR, "Elaboration Error: Too deep a hierarchy (for now)!");
            for (const std::string& sig : it->second.signals) signal(sig, inst_ns);  // elab.rs:218-239
            for (const CompSpec& sub : it->second.comps) instance(sub, inst_ns, false);
            path.pop_back();
            break;
        }
        }
    }
    void run() {  // elab.rs:244-265
        Ns ns;
        ns[""] = S21_GND;
        for (const std::string& s : ck.signals) signal(s, ns);
        for (const CompSpec& c : ck.comps) instance(c, ns, true);
        e.name_ptrs.resize(e.names.size());
        for (size_t k = 0; k < e.names.size(); k++) e.name_ptrs[k] = e.names[k].c_str();
    }
};

static void view(const s21_elab& e, s21_netlist& n) {
    n.n_vars = (int32_t)e.names.size();
    n.var_kind = e.kinds.data();
    n.var_name = e.name_ptrs.data();
    n.n_comps = (int32_t)e.comps.size();
    n.comps = e.comps.data();
    n.n_mos1_models = (int32_t)e.mos1_models.size();
    n.mos1_models = e.mos1_models.data();
    n.n_mos1_insts = (int32_t)e.mos1_insts.size();
    n.mos1_insts = e.mos1_insts.data();
    n.n_diode_models = (int32_t)e.diode_models.size();
    n.diode_models = e.diode_models.data();
    n.n_diode_insts = (int32_t)e.diode_insts.size();
    n.diode_insts = e.diode_insts.data();
    n.n_bsim4_models = (int32_t)e.bsim4_models.size();
    n.bsim4_models = e.bsim4_models.data();
    n.n_bsim4_insts = (int32_t)e.bsim4_insts.size();
    n.bsim4_insts = e.bsim4_insts.data();
}

static bool has_bsim4(const s21_netlist& nl) {
    for (int k = 0; k < nl.n_comps; k++)
        if (nl.comps[k].type == S21_COMP_BSIM4) return true;
    return false;
}

// SimOptions (protos/spice21.proto:136-142) -> Options (analysis.rs:665-688)
static s21_options sim_options(Reader r) {
    s21_options o;
    s21_default_options(&o);
    int f, wt;
    while (r.next(f, wt)) {
        if (wt != 2 || f < 1 || f > 5) {
            r.skip(wt);
            continue;
        }
        double v = wrapped_double(r.sub());
        if (f == 1) o.temp = v;
        else if (f == 2) o.tnom = v;
        else if (f == 3) o.gmin = v;
        else if (f == 4) o.iabstol = v;
        else o.reltol = v;
    }
    return o;
}

struct Call {
    bool has_ckt = false;
    s21_ckt ckt;
    s21_options opts;
    bool has_args = false;
    double tstop = 0.0, tstep = 0.0;
    std::vector<std::pair<std::string, double>> ic;
    uint64_t fstart = 0, fstop = 0, npts = 0;
};

// Op / Tran / Ac: ckt 1, opts 2, args 3.
static Call decode_call(const uint8_t* buf, size_t len, int kind) {
    Call c;
    s21_default_options(&c.opts);
    Reader r(buf, len);
    int f, wt;
    while (r.next(f, wt)) {
        if (f == 1 && wt == 2) {
            c.has_ckt = true;
            decode_circuit(r.sub(), c.ckt);
        } else if (f == 2 && wt == 2) {
            c.opts = sim_options(r.sub());
        } else if (f == 3 && wt == 2 && kind != 0) {
            c.has_args = true;
            Reader a = r.sub();
            int g, w;
            while (a.next(g, w)) {
                if (kind == 1 && g == 1 && w == 1)
                    c.tstop = a.f64();
                else if (kind == 1 && g == 2 && w == 1)
                    c.tstep = a.f64();
                else if (kind == 1 && g == 3 && w == 2) {
                    std::string k;
                    double v = 0.0;
                    double_map_entry(a.sub(), k, v);
                    c.ic.push_back({k, v});
                } else if (kind == 2 && g >= 1 && g <= 3 && w == 0) {
                    uint64_t v = a.varint();
                    (g == 1 ? c.fstart : g == 2 ? c.fstop : c.npts) = v;
                } else
                    a.skip(w);
            }
        } else
            r.skip(wt);
    }
    if (!c.has_ckt) fail(S21_ERR, "No Circuit Provided");  // proto.rs:60-64
    return c;
}

static uint8_t* to_heap(const std::vector<uint8_t>& v, size_t* out_len) {
    uint8_t* p = (uint8_t*)malloc(v.empty() ? 1 : v.size());
    if (!p) fail(S21_ERR, "out of memory");
    if (!v.empty()) memcpy(p, v.data(), v.size());
    *out_len = v.size();
    return p;
}

template <typename F>
static int guard_c(F&& f) {
    try {
        f();
        return S21_OK;
    } catch (Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (std::exception& e) {
        set_last_error(e.what());
        return S21_ERR;
    }
}

static void check_status(s21_batch* b) {
    int32_t st = 0;
    int64_t it = 0;
    if (s21_batch_get_status(b, 0, &st, &it) != S21_OK) fail(S21_ERR, "status unavailable");
    static const char* txt[] = {"", "Error", "Convergence Failed", "Singular Matrix", "Pivot Search Fail",
                                "Decode Error", "Invalid Argument", "Device Error", "Convergence Failed"};
    if (st != S21_OK) fail(st, st >= 0 && st <= 8 ? txt[st] : "Error");
}

struct BatchPtr {
    s21_batch* b = nullptr;
    ~BatchPtr() {
        if (b) s21_batch_free(b);
    }
};
struct LargePtr {
    s21_large* l = nullptr;
    ~LargePtr() {
        if (l) s21_large_free(l);
    }
};
static void check_status(s21_large* l) {
    int32_t st = 0;
    int64_t it = 0;
    if (s21_large_get_status(l, &st, &it) != S21_OK) fail(S21_ERR, "status unavailable");
    static const char* txt[] = {"", "Error", "Convergence Failed", "Singular Matrix", "Pivot Search Fail",
                                "Decode Error", "Invalid Argument", "Device Error", "Convergence Failed"};
    if (st != S21_OK) fail(st, st >= 0 && st <= 8 ? txt[st] : "Error");
}

static void engine_check(int rc) {
    if (rc != S21_OK) {
        char buf[1024];
        s21_last_error(buf, sizeof buf);
        fail(rc, buf);
    }
}

// Whether a batch call failed only because the circuit does not fit one workgroup's LDS (or the batched
// kernel's 16-bit record indices): such a circuit runs on the single-instance path instead.
static bool too_large_for_batch(int rc) {
    if (rc != S21_INVALID) return false;
    char buf[1024];
    s21_last_error(buf, sizeof buf);
    return std::strstr(buf, "circuit too large for the batched kernel") != nullptr;
}

}  // namespace s21

using namespace s21;

extern "C" {

int s21_ckt_decode(const uint8_t* buf, size_t len, s21_ckt** out) {
    return guard_c([&] {
        if (!out || (!buf && len)) fail(S21_INVALID, "bad arguments");
        std::unique_ptr<s21_ckt> c(new s21_ckt());
        decode_circuit(Reader(buf, len), *c);
        *out = c.release();
    });
}

int s21_ckt_add_mos0(s21_ckt* ckt, const char* name, int32_t mos_type) {
    return guard_c([&] {
        if (!ckt || !name) fail(S21_INVALID, "bad arguments");
        ckt->mos0[name] = mos_type == 1 ? 1 : 0;
    });
}

void s21_ckt_free(s21_ckt* ckt) { delete ckt; }

int s21_elaborate(const s21_ckt* ckt, s21_elab** out) {
    return guard_c([&] {
        if (!ckt || !out) fail(S21_INVALID, "bad arguments");
        std::unique_ptr<s21_elab> e(new s21_elab());
        Elaborator el(*ckt, *e);
        el.run();
        *out = e.release();
    });
}

int s21_elab_netlist(const s21_elab* e, s21_netlist* out) {
    return guard_c([&] {
        if (!e || !out) fail(S21_INVALID, "bad arguments");
        view(*e, *out);
    });
}

void s21_elab_free(s21_elab* e) { delete e; }

void s21_free_bytes(uint8_t* p) { free(p); }

// Op -> OpResult {map<string, double> vals = 1} (proto.rs:53-74)
int s21_op_bytes(const uint8_t* in, size_t len, uint8_t** out, size_t* out_len) {
    return guard_c([&] {
        if (!out || !out_len) fail(S21_INVALID, "bad arguments");
        Call c = decode_call(in, len, 0);
        s21_elab e;
        Elaborator(c.ckt, e).run();
        s21_netlist nl;
        view(e, nl);
        std::vector<double> x(nl.n_vars);
        // BSIM4 circuits, and circuits too large for the batched kernel, run on the single-instance path
        bool large = has_bsim4(nl);
        if (!large) {
            s21_group g{&nl, 1, 0, nullptr};
            BatchPtr bp;
            int rc = s21_batch_create(1, &g, &c.opts, S21_FP64, &bp.b);
            if (rc == S21_OK) rc = s21_batch_op(bp.b);
            if (too_large_for_batch(rc)) {
                large = true;
            } else {
                engine_check(rc);
                check_status(bp.b);
                engine_check(s21_batch_get_x(bp.b, 0, x.data()));
            }
        }
        if (large) {
            LargePtr lp;
            engine_check(s21_large_create(&nl, &c.opts, S21_FP64, &lp.l));
            engine_check(s21_large_op(lp.l));
            check_status(lp.l);
            engine_check(s21_large_get_x(lp.l, x.data()));
        }
        std::map<std::string, double> vals;  // OpResult::from keeps the last value of a repeated name
        for (int k = 0; k < nl.n_vars; k++) vals[e.names[k]] = x[k];
        Writer w;
        for (auto& kv : vals) {
            Writer ent;
            ent.str(1, kv.first);
            ent.f64(2, kv.second);
            w.bytes(1, ent.b);
        }
        *out = to_heap(w.b, out_len);
    });
}

// Tran -> TranResult {DoubleArray time = 1; map<string, DoubleArray> vals = 2} (proto.rs:76-110)
int s21_tran_bytes(const uint8_t* in, size_t len, uint8_t** out, size_t* out_len) {
    return guard_c([&] {
        if (!out || !out_len) fail(S21_INVALID, "bad arguments");
        Call c = decode_call(in, len, 1);
        if (c.tstep <= 0.0 && c.tstop > 0.0) fail(S21_INVALID, "tstep must be positive");
        s21_elab e;
        Elaborator el(c.ckt, e);
        el.run();
        s21_netlist nl;
        view(e, nl);
        std::vector<int32_t> ic_var;
        std::vector<double> ic_val;
        std::vector<const char*> ic_name;
        for (auto& kv : c.ic) {  // Tran::ic resolves the node by name (analysis.rs:510-525)
            if (kv.first.empty()) fail(S21_INVALID, "initial condition on ground");
            auto it = el.first.find(kv.first);
            if (it == el.first.end()) fail(S21_INVALID, "initial condition on unknown node: " + kv.first);
            ic_var.push_back(it->second);
            ic_val.push_back(kv.second);
            ic_name.push_back(kv.first.c_str());
        }
        s21_tran_options to{c.tstep, c.tstop, (int32_t)ic_var.size(), ic_var.data(), ic_val.data(), ic_name.data()};
        BatchPtr bp;
        LargePtr lp;
        int64_t np;
        int nv;
        std::vector<double> t, wv;
        std::vector<std::string> vnames;
        // BSIM4 circuits, and circuits too large for the batched kernel, run on the single-instance path
        bool large = has_bsim4(nl);
        if (!large) {
            s21_group g{&nl, 1, 0, nullptr};
            int rc = s21_batch_create(1, &g, &c.opts, S21_FP64, &bp.b);
            if (rc == S21_OK) rc = s21_batch_tran(bp.b, &to);
            large = too_large_for_batch(rc);
            if (!large) engine_check(rc);
        }
        if (large) {
            engine_check(s21_large_create(&nl, &c.opts, S21_FP64, &lp.l));
            engine_check(s21_large_tran(lp.l, &to));
            check_status(lp.l);
            np = s21_large_n_points(lp.l);
            nv = s21_large_n_vars(lp.l);
            t.resize(np);
            wv.resize((size_t)np * nv);
            std::vector<int32_t> all(nv);
            for (int k = 0; k < nv; k++) all[k] = k;
            engine_check(s21_large_get_axis(lp.l, t.data()));
            engine_check(s21_large_get_wave(lp.l, nv, all.data(), wv.data()));
            for (int k = 0; k < nv; k++) vnames.push_back(s21_large_var_name(lp.l, k));
        } else {
            check_status(bp.b);
            np = s21_batch_n_points(bp.b);
            nv = s21_batch_n_vars(bp.b, 0);
            t.resize(np);
            wv.resize((size_t)np * nv);
            engine_check(s21_batch_get_axis(bp.b, t.data()));
            engine_check(s21_batch_get_wave(bp.b, 0, 0, wv.data()));
            for (int k = 0; k < nv; k++) vnames.push_back(s21_batch_var_name(bp.b, 0, k));
        }
        Writer w;
        {
            Writer arr;
            arr.packed_f64(1, t.data(), t.size());
            w.bytes(1, arr.b);
        }
        std::map<std::string, std::vector<double>> vals;  // TranResult::end keys signals by name
        vals["time"] = t;
        for (int k = 0; k < nv; k++) {
            std::vector<double> col(np);
            for (int64_t p = 0; p < np; p++) col[p] = wv[(size_t)p * nv + k];
            vals[vnames[k]] = col;
        }
        for (auto& kv : vals) {
            Writer arr, ent;
            arr.packed_f64(1, kv.second.data(), kv.second.size());
            ent.str(1, kv.first);
            ent.bytes(2, arr.b);
            w.bytes(2, ent.b);
        }
        *out = to_heap(w.b, out_len);
    });
}

// Ac -> AcResult {DoubleArray freq = 1; map<string, ComplexArray> vals = 2} (proto.rs:112-145)
int s21_ac_bytes(const uint8_t* in, size_t len, uint8_t** out, size_t* out_len) {
    return guard_c([&] {
        if (!out || !out_len) fail(S21_INVALID, "bad arguments");
        Call c = decode_call(in, len, 2);
        s21_elab e;
        Elaborator(c.ckt, e).run();
        s21_netlist nl;
        view(e, nl);
        // the reference's BSIM4 has no AC stamps: Component::load_ac's default (comps/mod.rs:86-88)
        if (has_bsim4(nl)) fail(S21_ERR, "AC Not Implemented For This Component!");
        s21_ac_options ao{c.fstart, c.fstop, c.npts};
        s21_group g{&nl, 1, 0, nullptr};
        BatchPtr bp;
        engine_check(s21_batch_create(1, &g, &c.opts, S21_FP64, &bp.b));
        engine_check(s21_batch_ac(bp.b, &ao));
        check_status(bp.b);
        // the reference's `ac` writes its result files into the working directory (analysis.rs:783-827)
        engine_check(s21_batch_write_ac_results(bp.b, 0, 0, "."));
        const int64_t np = s21_batch_n_points(bp.b);
        const int nv = s21_batch_n_vars(bp.b, 0);
        std::vector<double> f(np), wv((size_t)np * nv * 2);
        engine_check(s21_batch_get_axis(bp.b, f.data()));
        engine_check(s21_batch_get_wave(bp.b, 0, 0, wv.data()));
        Writer w;
        {
            Writer arr;
            arr.packed_f64(1, f.data(), f.size());
            w.bytes(1, arr.b);
        }
        std::map<std::string, int> idx;
        for (int k = 0; k < nv; k++) idx[s21_batch_var_name(bp.b, 0, k)] = k;
        for (auto& kv : idx) {
            Writer arr, ent;
            for (int64_t p = 0; p < np; p++) {
                Writer cn;
                cn.f64(1, wv[((size_t)p * nv + kv.second) * 2 + 0]);
                cn.f64(2, wv[((size_t)p * nv + kv.second) * 2 + 1]);
                arr.bytes(1, cn.b);
            }
            ent.str(1, kv.first);
            ent.bytes(2, arr.b);
            w.bytes(2, ent.b);
        }
        *out = to_heap(w.b, out_len);
    });
}

}  // extern "C"
