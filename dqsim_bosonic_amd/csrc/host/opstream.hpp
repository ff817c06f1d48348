// opstream.hpp -- decoder of the opcode-stream wire format (include/dqsim_opstream.h)
// into the same instruction list the JSON decoder produces (circuit.hpp).
#pragma once
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/dqsim_opstream.h"
#include "circuit.hpp"

namespace dqh {

static_assert(sizeof(dq_op_record) == 64, "dq_op_record is 64 bytes");

inline const char *opcode_kind(int op) {
    static const char *names[DQ_OP_COUNT] = {
        "id", "x", "y", "z", "h", "s", "sdg", "t", "tdg", "sx", "sxdg", "u0",
        "u3", "u", "u2", "u1", "p", "rx", "ry", "rz",
        "cx", "cz", "cy", "ch", "csx", "swap",
        "crx", "cry", "crz", "cu1", "cp", "cu3", "cu", "rxx", "rzz",
        "ccx", "cswap", "rccx", "rc3x", "c3x", "c3sqrtx", "c4x",
        "gate", "measure", "reset", "conditional", "barrier", "classical"};
    return (op >= 0 && op < DQ_OP_COUNT) ? names[op] : nullptr;
}

struct OpstreamError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// One instruction from records[i] (and the following records of a
// conditional); advances i.
inline Instr opstream_instr(const dq_op_record *rec, size_t count, size_t &i, const std::vector<std::string> &names) {
    if (i >= count) throw OpstreamError("conditional record at the end of the stream");
    const dq_op_record &r = rec[i++];
    const char *kind = opcode_kind(r.opcode);
    if (!kind) throw OpstreamError("unknown opcode " + std::to_string(r.opcode));
    Instr in;
    in.kind = kind;
    const auto &specs = kind_specs();
    auto it = specs.find(in.kind);
    if (it != specs.end()) {
        const size_t nq = it->second.qubits.size(), np = it->second.params.size();
        for (size_t j = 0; j < nq; ++j) in.qubits.push_back(r.qubits[j]);
        for (size_t j = 0; j < np; ++j) in.params.push_back(r.params[j]);
        return in;
    }
    switch (r.opcode) {
    case DQ_OP_GATE:
    case DQ_OP_CLASSICAL:
        if (r.cbit >= names.size()) throw OpstreamError("name index out of range");
        in.name = names[r.cbit];
        if (r.opcode == DQ_OP_GATE) {
            if (r.nq > 5) throw OpstreamError("a generic gate record holds at most 5 qubits");
            for (unsigned j = 0; j < r.nq; ++j) in.qubits.push_back(r.qubits[j]);
        }
        return in;
    case DQ_OP_MEASURE:
        in.qubits.push_back(r.qubits[0]);
        in.cbit = r.cbit;
        return in;
    case DQ_OP_RESET:
        in.qubits.push_back(r.qubits[0]);
        return in;
    case DQ_OP_CONDITIONAL: {
        in.creg_base = r.qubits[0];
        in.creg_size = r.qubits[1];
        uint64_t v;
        std::memcpy(&v, r.params, sizeof v);
        in.creg_value = v;
        in.op = std::make_shared<Instr>(opstream_instr(rec, count, i, names));
        return in;
    }
    default:  // barrier
        return in;
    }
}

inline Circuit decode_opstream(const dq_op_record *rec, size_t count, uint64_t n_qubits, uint64_t n_cbits,
                               const std::vector<std::string> &names) {
    Circuit c;
    c.qregs.push_back(Register{"q", n_qubits, 0});
    if (n_cbits) c.cregs.push_back(Register{"c", n_cbits, 0});
    c.instructions.reserve(count);
    for (size_t i = 0; i < count;) c.instructions.push_back(opstream_instr(rec, count, i, names));
    return c;
}

}  // namespace dqh
