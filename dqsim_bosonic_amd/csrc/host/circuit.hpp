// circuit.hpp -- decoder for the circuit wire format.
//
// Restates the serde schema of the reference's src/types.rs:7-135: an object
// with `qregs` (map name -> {name, size, base}), optional `cregs` (same), and
// `instructions`, each tagged by `kind` (snake_case variant names) with the
// variant's fields.  Unknown fields are ignored, as serde does by default;
// missing fields, wrong types and unknown kinds are errors.
#pragma once
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "json.hpp"

namespace dqh {

struct Instr {
    std::string kind;              // instruction kind (types.rs:60-135, snake_case)
    std::vector<uint64_t> qubits;  // gate qubits in matrix order (first = MSB)
    std::vector<double> params;    // gate parameters in field order
    std::string name;              // `gate` / `classical`
    uint64_t cbit = 0;             // `measure`
    uint64_t creg_base = 0, creg_size = 0, creg_value = 0;  // `conditional`
    std::shared_ptr<Instr> op;     // `conditional`
};

struct Register {
    std::string name;
    uint64_t size = 0, base = 0;
};

struct Circuit {
    std::vector<Register> qregs, cregs;
    std::vector<Instr> instructions;
    // types.rs:24-30 / :32-39
    static uint64_t extent(const std::vector<Register> &regs) {
        uint64_t m = 0;
        for (const auto &r : regs) m = std::max(m, r.base + r.size);
        return m;
    }
    uint64_t num_qubits() const { return extent(qregs); }
    uint64_t num_cbits() const { return extent(cregs); }
};

struct KindSpec {
    std::vector<const char *> qubits;
    std::vector<const char *> params;
};

// Field lists of every gate-like kind, in the reference's matrix order.
inline const std::unordered_map<std::string, KindSpec> &kind_specs() {
    static const std::unordered_map<std::string, KindSpec> m = [] {
        std::unordered_map<std::string, KindSpec> t;
        for (const char *k : {"id", "x", "y", "z", "h", "s", "sdg", "t", "tdg", "sx", "sxdg", "u0"})
            t[k] = {{"qubit"}, {}};
        t["u3"] = {{"qubit"}, {"theta", "phi", "lam"}};
        t["u"] = {{"qubit"}, {"theta", "phi", "lam"}};
        t["u2"] = {{"qubit"}, {"phi", "lam"}};
        t["u1"] = {{"qubit"}, {"lam"}};
        t["p"] = {{"qubit"}, {"lam"}};
        t["rx"] = {{"qubit"}, {"theta"}};
        t["ry"] = {{"qubit"}, {"theta"}};
        t["rz"] = {{"qubit"}, {"phi"}};
        for (const char *k : {"cx", "cz", "cy", "ch", "csx"}) t[k] = {{"control", "target"}, {}};
        t["swap"] = {{"a", "b"}, {}};
        t["crx"] = {{"control", "target"}, {"theta"}};
        t["cry"] = {{"control", "target"}, {"theta"}};
        t["crz"] = {{"control", "target"}, {"lam"}};
        t["cu1"] = {{"control", "target"}, {"lam"}};
        t["cp"] = {{"control", "target"}, {"lam"}};
        t["cu3"] = {{"control", "target"}, {"theta", "phi", "lam"}};
        t["cu"] = {{"control", "target"}, {"theta", "phi", "lam", "gamma"}};
        t["rxx"] = {{"a", "b"}, {"theta"}};
        t["rzz"] = {{"a", "b"}, {"theta"}};
        t["ccx"] = {{"control1", "control2", "target"}, {}};
        t["cswap"] = {{"control", "target1", "target2"}, {}};
        t["rccx"] = {{"control1", "control2", "target"}, {}};
        for (const char *k : {"rc3x", "c3x", "c3sqrtx"}) t[k] = {{"control1", "control2", "control3", "target"}, {}};
        t["c4x"] = {{"control1", "control2", "control3", "control4", "target"}, {}};
        return t;
    }();
    return m;
}

class Decoder {
   public:
    Circuit decode(const std::string &text) {
        Json root = JsonParser(text).parse();
        if (root.type != Json::Object) throw JsonError("invalid type: expected struct Circuit");
        Circuit c;
        const Json *q = root.get("qregs");
        if (!q) throw JsonError("missing field `qregs`");
        c.qregs = regs(*q);
        if (const Json *cr = root.get("cregs")) c.cregs = regs(*cr);
        const Json *ins = root.get("instructions");
        if (!ins) throw JsonError("missing field `instructions`");
        if (ins->type != Json::Array) throw JsonError("invalid type: expected a sequence for `instructions`");
        c.instructions.reserve(ins->arr.size());
        for (const Json &j : ins->arr) c.instructions.push_back(instr(j));
        return c;
    }

   private:
    static uint64_t as_usize(const Json *v, const char *field) {
        if (!v) throw JsonError(std::string("missing field `") + field + "`");
        if (v->type != Json::Number) throw JsonError(std::string("invalid type for `") + field + "`, expected usize");
        if (!v->is_integer() || v->text[0] == '-')
            throw JsonError(std::string("invalid value for `") + field + "`: " + v->text + ", expected usize");
        errno = 0;
        char *end = nullptr;
        const unsigned long long x = std::strtoull(v->text.c_str(), &end, 10);
        if (errno == ERANGE) throw JsonError(std::string("number out of range for `") + field + "`");
        return uint64_t(x);
    }
    static double as_f64(const Json *v, const char *field) {
        if (!v) throw JsonError(std::string("missing field `") + field + "`");
        if (v->type != Json::Number) throw JsonError(std::string("invalid type for `") + field + "`, expected f64");
        return std::strtod(v->text.c_str(), nullptr);
    }
    static std::string as_str(const Json *v, const char *field) {
        if (!v) throw JsonError(std::string("missing field `") + field + "`");
        if (v->type != Json::String) throw JsonError(std::string("invalid type for `") + field + "`, expected a string");
        return v->text;
    }
    static std::vector<Register> regs(const Json &m) {
        if (m.type != Json::Object) throw JsonError("invalid type: expected a map of registers");
        std::vector<Register> out;
        for (const auto &kv : m.obj) {
            if (kv.second.type != Json::Object) throw JsonError("invalid type: expected struct Register");
            Register r;
            r.name = as_str(kv.second.get("name"), "name");
            r.size = as_usize(kv.second.get("size"), "size");
            r.base = as_usize(kv.second.get("base"), "base");
            out.push_back(r);
        }
        return out;
    }
    Instr instr(const Json &j) {
        if (j.type != Json::Object) throw JsonError("invalid type: expected internally tagged enum Instruction");
        Instr in;
        in.kind = as_str(j.get("kind"), "kind");
        const auto &specs = kind_specs();
        auto it = specs.find(in.kind);
        if (it != specs.end()) {
            for (const char *f : it->second.qubits) in.qubits.push_back(as_usize(j.get(f), f));
            for (const char *f : it->second.params) in.params.push_back(as_f64(j.get(f), f));
            return in;
        }
        if (in.kind == "gate") {
            in.name = as_str(j.get("name"), "name");
            const Json *p = j.get("params");
            if (!p) throw JsonError("missing field `params`");
            if (p->type != Json::Array) throw JsonError("invalid type for `params`, expected a sequence");
            for (const Json &x : p->arr) in.params.push_back(as_f64(&x, "params"));
            const Json *qs = j.get("qubits");
            if (!qs) throw JsonError("missing field `qubits`");
            if (qs->type != Json::Array) throw JsonError("invalid type for `qubits`, expected a sequence");
            for (const Json &x : qs->arr) in.qubits.push_back(as_usize(&x, "qubits"));
            return in;
        }
        if (in.kind == "measure") {
            in.qubits.push_back(as_usize(j.get("qubit"), "qubit"));
            in.cbit = as_usize(j.get("cbit"), "cbit");
            return in;
        }
        if (in.kind == "reset") {
            in.qubits.push_back(as_usize(j.get("qubit"), "qubit"));
            return in;
        }
        if (in.kind == "conditional") {
            const Json *c = j.get("condition");
            if (!c) throw JsonError("missing field `condition`");
            if (c->type != Json::Object) throw JsonError("invalid type for `condition`, expected struct Condition");
            in.creg_base = as_usize(c->get("creg_base"), "creg_base");
            in.creg_size = as_usize(c->get("creg_size"), "creg_size");
            in.creg_value = as_usize(c->get("creg_value"), "creg_value");
            const Json *op = j.get("op");
            if (!op) throw JsonError("missing field `op`");
            in.op = std::make_shared<Instr>(instr(*op));
            return in;
        }
        if (in.kind == "barrier") return in;
        if (in.kind == "classical") {
            in.name = as_str(j.get("name"), "name");
            return in;
        }
        throw JsonError("unknown variant `" + in.kind + "`");
    }
};

}  // namespace dqh
